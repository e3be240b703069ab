// pdm_bonds.hip -- GPU family (bond) search, boundary-effect factors and per-entry coefficients.
//
// Replaces InitBond's O(N^2) pair loop (pdDataManager2.cpp:12-70) with a cell-binned search:
//   1. bin nodes into cubic cells of edge delta + 1.5 dx (the reference's prefilter radius),
//      counting sort by cell;
//   2. count pass: thread per node i tests every node j != i of the 27 neighbouring cells with
//      the reference predicate evaluated on (lower id, higher id), exactly as InitBond visits a
//      pair, so both ends of a pair get the same answer;
//   3. exclusive scan -> row_ptr; fill pass writes the neighbours; each row is sorted
//      ascending in LDS (InitBond appends neighbours in id order);
//   4. slit cuts, again per pair in (lower, higher) order.
// Then InitBdryEffectFactor (pdDataManager.cpp:1340-1473): full families -> 1, type-one band
// on the host (pow() of the reference), energy branch here, one thread per node summing its
// family in ascending order as CalcBdryEffFactor does (pdDataManager2.cpp:1353-1477).
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

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;   // elements per thread
constexpr int kScanTile = kScanBlock * kScanItems;

// Exclusive scan of one tile per block; block total to sums[blockIdx].
__global__ void k_scan_tile(const int64_t* __restrict__ in, int64_t* __restrict__ out, int64_t* __restrict__ sums,
                            int64_t n)
{
    __shared__ int64_t sh[kScanBlock];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t tsum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        int64_t idx = base + k;
        v[k] = (idx < n) ? in[idx] : 0;
        tsum += v[k];
    }
    sh[threadIdx.x] = tsum;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        int64_t add = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t run = sh[threadIdx.x] - tsum;  // exclusive prefix of this thread
    for (int k = 0; k < kScanItems; ++k) {
        int64_t idx = base + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanBlock - 1 && sums) sums[blockIdx.x] = sh[threadIdx.x];
}

__global__ void k_scan_add(int64_t* __restrict__ out, const int64_t* __restrict__ offs, int64_t n)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) out[idx] += offs[idx / kScanTile];
}

}  // namespace

// out[i] = sum_{k<i} in[k] for i in [0, n]; out has n+1 entries.
int scan_exclusive(pdm_handle* h, const int64_t* in, int64_t* out, int64_t n)
{
    hipStream_t s = h->stream;
    int64_t nb = (n + 1 + kScanTile - 1) / kScanTile;
    int64_t* sums = nullptr;
    int64_t* sums_scan = nullptr;
    // scan n+1 elements with in[n] treated as 0: copy through a padded buffer
    int64_t* tmp = nullptr;
    DevTemps temps(s);
    temps.add(&tmp);
    temps.add(&sums);
    temps.add(&sums_scan);
    HIPCHK(hipMalloc(&tmp, sizeof(int64_t) * (n + 1)));
    HIPCHK(hipMemsetAsync(tmp + n, 0, sizeof(int64_t), s));
    if (n) HIPCHK(hipMemcpyAsync(tmp, in, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMalloc(&sums, sizeof(int64_t) * nb));
    HIPCHK(hipMalloc(&sums_scan, sizeof(int64_t) * (nb + 1)));
    k_scan_tile<<<dim3((unsigned)nb), dim3(kScanBlock), 0, s>>>(tmp, out, sums, n + 1);
    HIPCHK(hipGetLastError());
    if (nb > 1) {
        int rc = scan_exclusive(h, sums, sums_scan, nb);
        if (rc) return rc;
        int64_t m = n + 1;
        k_scan_add<<<dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s>>>(out, sums_scan, m);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    return PDM_OK;
}

namespace {

__device__ __forceinline__ int64_t cell_of(const BondArgs& a, double x, int c)
{
    int64_t q = (int64_t)floor((x - a.lo[c]) / a.h);
    if (q < 0) q = 0;
    if (q >= a.nc[c]) q = a.nc[c] - 1;
    return q;
}

__global__ void k_bin_count(BondArgs a, int64_t* __restrict__ cell_id, int64_t* __restrict__ cnt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    d4 x = a.X[i];
    int64_t q0 = cell_of(a, x.x, 0), q1 = cell_of(a, x.y, 1), q2 = cell_of(a, x.z, 2);
    int64_t c = q0 + a.nc[0] * (q1 + a.nc[1] * q2);
    cell_id[i] = c;
    atomicAdd((unsigned long long*)&cnt[c], 1ULL);
}

__global__ void k_bin_scatter(BondArgs a, const int64_t* __restrict__ cell_id, int64_t* __restrict__ cursor,
                              int32_t* __restrict__ cell_nodes)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    unsigned long long slot = atomicAdd((unsigned long long*)&cursor[cell_id[i]], 1ULL);
    cell_nodes[slot] = (int32_t)i;
}

// Count (FILL=false) or write (FILL=true) node i's family.
template <bool FILL>
__global__ void k_family(BondArgs a, const int64_t* __restrict__ cell_id, const int64_t* __restrict__ cell_start,
                         const int32_t* __restrict__ cell_nodes, int64_t* __restrict__ count,
                         const int64_t* __restrict__ row_ptr, int32_t* __restrict__ col,
                         unsigned long long* __restrict__ min_bits)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    d4 xi = a.X[i];
    int64_t c0 = cell_id[i];
    int64_t q0 = c0 % a.nc[0], q1 = (c0 / a.nc[0]) % a.nc[1], q2 = c0 / (a.nc[0] * a.nc[1]);
    int64_t k = 0;
    int64_t out = FILL ? row_ptr[i] : 0;
    double dmin_local = a.horizon;  // InitBond starts _minBondLength at the horizon
    for (int64_t z = q2 - 1; z <= q2 + 1; ++z) {
        if (z < 0 || z >= a.nc[2]) continue;
        for (int64_t y = q1 - 1; y <= q1 + 1; ++y) {
            if (y < 0 || y >= a.nc[1]) continue;
            for (int64_t x = q0 - 1; x <= q0 + 1; ++x) {
                if (x < 0 || x >= a.nc[0]) continue;
                int64_t cc = x + a.nc[0] * (y + a.nc[1] * z);
                for (int64_t s = cell_start[cc]; s < cell_start[cc + 1]; ++s) {
                    int32_t j = cell_nodes[s];
                    if (j == i) continue;
                    d4 xj = a.X[j];
                    double dist;
                    bool ok;
                    if (i < j) ok = g_bonded(xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, a.dx, a.horizon, &dist);
                    else ok = g_bonded(xj.x, xj.y, xj.z, xi.x, xi.y, xi.z, a.dx, a.horizon, &dist);
                    if (!ok) continue;
                    if (FILL) col[out + k] = j;
                    else if (i < j) dmin_local = (dist < dmin_local) ? dist : dmin_local;
                    ++k;
                }
            }
        }
    }
    if (!FILL) {
        count[i] = k;
        atomicMin(min_bits, (unsigned long long)__double_as_longlong(dmin_local));
    }
}

// Sort each row ascending (bitonic in LDS), rows up to kMaxFamily entries.
__global__ void k_sort_rows(const int64_t* __restrict__ row_ptr, int32_t* __restrict__ col, int64_t n)
{
    __shared__ int32_t sh[kMaxFamily];
    int64_t i = blockIdx.x;
    if (i >= n) return;
    int64_t b = row_ptr[i];
    int len = (int)(row_ptr[i + 1] - b);
    int P = 1;
    while (P < len) P <<= 1;
    for (int t = threadIdx.x; t < P; t += blockDim.x) sh[t] = (t < len) ? col[b + t] : INT32_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int t = threadIdx.x; t < P; t += blockDim.x) {
                int p = t ^ jj;
                if (p > t) {
                    bool up = ((t & k) == 0);
                    int32_t va = sh[t], vb = sh[p];
                    if ((va > vb) == up) {
                        sh[t] = vb;
                        sh[p] = va;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < len; t += blockDim.x) col[b + t] = sh[t];
}

// Slit cuts per entry, pair evaluated as (lower id, higher id).
__global__ void k_cuts(BondArgs a, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                       uint8_t* __restrict__ cut)
{
    int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    int lane = threadIdx.x % 64;
    if (i >= a.n) return;
    d4 xi = a.X[i];
    for (int64_t e = row_ptr[i] + lane; e < row_ptr[i + 1]; e += 64) {
        int32_t j = col[e];
        d4 xj = a.X[j];
        double pa[3], pb[3];
        if (i < j) { pa[0] = xi.x; pa[1] = xi.y; pa[2] = xi.z; pb[0] = xj.x; pb[1] = xj.y; pb[2] = xj.z; }
        else { pa[0] = xj.x; pa[1] = xj.y; pa[2] = xj.z; pb[0] = xi.x; pb[1] = xi.y; pb[2] = xi.z; }
        uint8_t c = 0;
        for (int s = 0; s < a.nslit; ++s)
            if (g_slit_cut(a.slits[s].type, a.slits[s].axis, a.slits[s].p, a.slits[s].w, pa, pb)) c = 1;
        cut[e] = c;
    }
}

struct BdryArgs {
    const d4* X;
    const int64_t* row_ptr;
    const int32_t* col;
    const uint8_t* cut;
    const int32_t* mat;
    const int8_t* branch;   // 0 full family, 1 type-one (host value), 2 energy branch
    double* be;
    int64_t n;
    double horizon, dx, vol;
    bool original_cci;
    double spring[kMaxMat], ecrit[kMaxMat], E[kMaxMat];
};

// CalcBdryEffFactor (pdDataManager2.cpp:1353-1477), same operation order, one thread per node.
__global__ void k_bdryeff_energy(BdryArgs a)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n || a.branch[i] != 2) return;
    const double e11 = 1.0e-4;
    int m = a.mat[i];
    double ecrit = a.ecrit[m], spring = a.spring[m];
    double eps11 = e11, eps22 = e11, eps33 = e11;
    double eps12 = 0.0, eps13 = 0.0, eps21 = 0.0, eps23 = 0.0, eps31 = 0.0, eps32 = 0.0;
    d4 xi = a.X[i];
    double sum = 0.0;
    // slit cuts do not enter: the factor is that of the intact family (a slit is a
    // pre-crack introduced after the surface calibration)
    for (int64_t e = a.row_ptr[i]; e < a.row_ptr[i + 1]; ++e) {
        d4 xj = a.X[a.col[e]];
        double xi1 = xj.x - xi.x, xi2 = xj.y - xi.y, xi3 = xj.z - xi.z;
        double eta1 = eps11 * xi1 + eps12 * xi2 + eps13 * xi3;
        double eta2 = eps21 * xi1 + eps22 * xi2 + eps23 * xi3;
        double eta3 = eps31 * xi1 + eps32 * xi2 + eps33 * xi3;
        double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
        double p1 = xi1 + eta1, p2 = xi2 + eta2, p3 = xi3 + eta3;
        double p = sqrt(p1 * p1 + p2 * p2 + p3 * p3);
        double u = p - r;
        double st = u / r;
        double el = (u <= 0 || (u > 0 && st <= ecrit)) ? 0.5 * (spring / r) * u * u : 0.0;
        double beta = g_beta(r, a.horizon, a.dx, a.original_cci);
        double elj = el * beta * a.vol;
        sum += 0.5 * elj;
    }
    double p_inf = 3.0 * a.E[m] * e11 * e11;
    a.be[i] = p_inf / sum;
}

}  // namespace

int build_bonds(pdm_handle* h)
{
    const Deck& d = h->deck;
    hipStream_t s = h->stream;
    int64_t n = h->n;
    BondArgs a{};
    a.n = n;
    a.dx = d.dx;
    a.horizon = d.horizon;
    a.tol = d.tol;
    a.h = d.horizon + 1.5 * d.dx;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c) { lo[c] = 1e300; hi[c] = -1e300; }
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], h->X[3 * i + c]);
            hi[c] = std::max(hi[c], h->X[3 * i + c]);
        }
    for (int c = 0; c < 3; ++c) {
        if (n == 0) { lo[c] = 0; hi[c] = 0; }
        a.lo[c] = lo[c];
        a.nc[c] = (int64_t)std::floor((hi[c] - lo[c]) / a.h) + 1;
    }
    if ((double)a.nc[0] * (double)a.nc[1] * (double)a.nc[2] > 4.0e9) {
        h->err = "bounding box too large for the cell grid";
        return PDM_E_LIMIT;
    }
    // slits: frames computed once here, as the oracle computes them per pair
    a.nslit = (int)d.slits.size();
    if (a.nslit > kMaxSlit) {
        h->err = "too many slits";
        return PDM_E_LIMIT;
    }
    for (int k = 0; k < a.nslit; ++k) {
        const Slit& sl = d.slits[k];
        SlitDev& o = a.slits[k];
        std::memset(&o, 0, sizeof(o));
        o.type = sl.type;
        o.axis = -1;
        for (int q = 0; q < 11; ++q) o.p[q] = sl.p[q];
        if (sl.type == 1) {
            for (int c = 0; c < 3; ++c)
                if (sl.p[2 * c] == sl.p[2 * c + 1]) { o.axis = c; break; }
        } else {
            double nn = std::sqrt(sl.p[3] * sl.p[3] + sl.p[4] * sl.p[4] + sl.p[5] * sl.p[5]);
            double nv[3] = {sl.p[3] / nn, sl.p[4] / nn, sl.p[5] / nn};
            double tn = sl.p[6] * nv[0] + sl.p[7] * nv[1] + sl.p[8] * nv[2];
            double tt[3] = {sl.p[6] - tn * nv[0], sl.p[7] - tn * nv[1], sl.p[8] - tn * nv[2]};
            double tl = std::sqrt(tt[0] * tt[0] + tt[1] * tt[1] + tt[2] * tt[2]);
            double tv[3] = {tt[0] / tl, tt[1] / tl, tt[2] / tl};
            for (int c = 0; c < 3; ++c) { o.p[3 + c] = nv[c]; o.p[6 + c] = tv[c]; }
            o.w[0] = nv[1] * tv[2] - nv[2] * tv[1];
            o.w[1] = nv[2] * tv[0] - nv[0] * tv[2];
            o.w[2] = nv[0] * tv[1] - nv[1] * tv[0];
        }
    }

    // positions
    std::vector<d4> Xh(n);
    for (int64_t i = 0; i < n; ++i) Xh[i] = d4{h->X[3 * i], h->X[3 * i + 1], h->X[3 * i + 2], 0.0};
    HIPCHK(hipMalloc(&h->Xd, sizeof(d4) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMemcpyAsync(h->Xd, Xh.data(), sizeof(d4) * n, hipMemcpyHostToDevice, s));
    a.X = h->Xd;

    int64_t ncell = a.nc[0] * a.nc[1] * a.nc[2];
    int64_t *cell_id = nullptr, *cell_cnt = nullptr, *cell_start = nullptr, *cursor = nullptr, *cnt = nullptr;
    int32_t* cell_nodes = nullptr;
    unsigned long long* min_bits = nullptr;
    int8_t* branch_d = nullptr;
    DevTemps temps(s);  // the cell-list scratch, released on every return path
    temps.add(&cell_id);
    temps.add(&cell_cnt);
    temps.add(&cell_start);
    temps.add(&cursor);
    temps.add(&cnt);
    temps.add(&cell_nodes);
    temps.add(&min_bits);
    temps.add(&branch_d);
    HIPCHK(hipMalloc(&cell_id, sizeof(int64_t) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMalloc(&cell_cnt, sizeof(int64_t) * ncell));
    HIPCHK(hipMalloc(&cell_start, sizeof(int64_t) * (ncell + 1)));
    HIPCHK(hipMalloc(&cursor, sizeof(int64_t) * (ncell + 1)));
    HIPCHK(hipMalloc(&cell_nodes, sizeof(int32_t) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMalloc(&cnt, sizeof(int64_t) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMalloc(&min_bits, sizeof(unsigned long long)));
    HIPCHK(hipMemsetAsync(cell_cnt, 0, sizeof(int64_t) * ncell, s));
    double hz = d.horizon;
    unsigned long long hb;
    std::memcpy(&hb, &hz, sizeof hb);
    HIPCHK(hipMemcpyAsync(min_bits, &hb, sizeof hb, hipMemcpyHostToDevice, s));
    unsigned nb256 = (unsigned)((n + 255) / 256);
    if (n) {
        k_bin_count<<<dim3(nb256), dim3(256), 0, s>>>(a, cell_id, cell_cnt);
        HIPCHK(hipGetLastError());
    }
    int rc = scan_exclusive(h, cell_cnt, cell_start, ncell);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(cursor, cell_start, sizeof(int64_t) * (ncell + 1), hipMemcpyDeviceToDevice, s));
    if (n) {
        k_bin_scatter<<<dim3(nb256), dim3(256), 0, s>>>(a, cell_id, cursor, cell_nodes);
        HIPCHK(hipGetLastError());
        k_family<false><<<dim3(nb256), dim3(256), 0, s>>>(a, cell_id, cell_start, cell_nodes, cnt, nullptr,
                                                           nullptr, min_bits);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMalloc(&h->rp, sizeof(int64_t) * (n + 1)));
    rc = scan_exclusive(h, cnt, h->rp, n);
    if (rc) return rc;
    int64_t E = 0;
    HIPCHK(hipMemcpyAsync(&E, h->rp + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hb, min_bits, sizeof hb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memcpy(&h->min_bond, &hb, sizeof hb);
    h->E = E;
    // row lengths for the sort limit and the boundary-effect branches
    std::vector<int64_t> rph(n + 1);
    HIPCHK(hipMemcpy(rph.data(), h->rp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    int64_t maxrow = 0;
    for (int64_t i = 0; i < n; ++i) maxrow = std::max(maxrow, rph[i + 1] - rph[i]);
    if (maxrow > kMaxFamily) {
        h->err = "a family exceeds " + std::to_string(kMaxFamily) + " entries";
        return PDM_E_LIMIT;
    }
    HIPCHK(hipMalloc(&h->col, sizeof(int32_t) * std::max<int64_t>(E, 1)));
    HIPCHK(hipMalloc(&h->cut, std::max<int64_t>(E, 1)));
    if (n) {
        k_family<true><<<dim3(nb256), dim3(256), 0, s>>>(a, cell_id, cell_start, cell_nodes, nullptr, h->rp, h->col,
                                                          nullptr);
        HIPCHK(hipGetLastError());
        k_sort_rows<<<dim3((unsigned)n), dim3(256), 0, s>>>(h->rp, h->col, n);
        HIPCHK(hipGetLastError());
        unsigned nbw = (unsigned)((n + 3) / 4);
        k_cuts<<<dim3(nbw), dim3(256), 0, s>>>(a, h->rp, h->col, h->cut);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));

    // InitBdryEffectFactor branches (pdDataManager.cpp:1340-1473)
    h->max_family = count_max_family(d.horizon, d.dx, d.tol);
    std::vector<int8_t> branch(n, 0);
    std::vector<double> beh(n, 1.0);
    if (d.bdryeff_on) {
        double x1L = -d.nx * 0.5 * d.dx + d.cx, x1H = d.nx * 0.5 * d.dx + d.cx;
        double x2L = -d.ny * 0.5 * d.dx + d.cy, x2H = d.ny * 0.5 * d.dx + d.cy;
        double x3L = -d.nz * 0.5 * d.dx + d.cz, x3H = d.nz * 0.5 * d.dx + d.cz;
        double ns = ((int)(d.horizon / d.dx)) * d.dx;
        for (int64_t i = 0; i < n; ++i) {
            double x1 = h->X[3 * i], x2 = h->X[3 * i + 1], x3 = h->X[3 * i + 2];
            int m = h->mat[i];
            double sp = h->spring[m], E0 = d.mats[m].E;
            if (rph[i + 1] - rph[i] == h->max_family) continue;
            branch[i] = 1;
            if (x1 >= x1H - ns && x1 <= x1H && x2 >= x2L + ns && x2 <= x2H - ns && x3 >= x3L + ns && x3 <= x3H - ns)
                beh[i] = bdryeff_type_one(d, sp, E0, x1H - x1);
            else if (x1 >= x1L && x1 <= x1L + ns && x2 >= x2L + ns && x2 <= x2H - ns && x3 >= x3L + ns &&
                     x3 <= x3H - ns)
                beh[i] = bdryeff_type_one(d, sp, E0, x1 - x1L);
            else if (x2 >= x2H - ns && x2 <= x2H && x1 >= x1L + ns && x1 <= x1H - ns && x3 >= x3L + ns &&
                     x3 <= x3H - ns)
                beh[i] = bdryeff_type_one(d, sp, E0, x2H - x2);
            else if (x2 >= x2L && x2 <= x2L + ns && x1 >= x1L + ns && x1 <= x1H - ns && x3 >= x3L + ns &&
                     x3 <= x3H - ns)
                beh[i] = bdryeff_type_one(d, sp, E0, x2 - x2L);
            else if (x3 >= x3H - ns && x3 <= x3H && x1 >= x1L + ns && x1 <= x1H - ns && x2 >= x2L + ns &&
                     x2 <= x2H - ns)
                beh[i] = bdryeff_type_one(d, sp, E0, x3H - x3);
            else if (x3 >= x3L && x3 <= x3L + ns && x1 >= x1L + ns && x1 <= x1H - ns && x2 >= x2L + ns &&
                     x2 <= x2H - ns)
                beh[i] = bdryeff_type_one(d, sp, E0, x3 - x3L);
            else
                branch[i] = 2;
        }
    }
    HIPCHK(hipMalloc(&h->be, sizeof(double) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMalloc(&branch_d, std::max<int64_t>(n, 1)));
    HIPCHK(hipMemcpyAsync(h->be, beh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(branch_d, branch.data(), n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMalloc(&h->matd, sizeof(int32_t) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMemcpyAsync(h->matd, h->mat.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    BdryArgs ba{};
    ba.X = h->Xd;
    ba.row_ptr = h->rp;
    ba.col = h->col;
    ba.cut = h->cut;
    ba.mat = h->matd;
    ba.branch = branch_d;
    ba.be = h->be;
    ba.n = n;
    ba.horizon = d.horizon;
    ba.dx = d.dx;
    ba.vol = std::pow(d.dx, 3.0);
    ba.original_cci = d.cci_on && !d.cci_mdf;
    for (size_t m = 0; m < d.mats.size() && m < (size_t)kMaxMat; ++m) {
        ba.spring[m] = h->spring[m];
        ba.ecrit[m] = d.mats[m].ecrit;
        ba.E[m] = d.mats[m].E;
    }
    if (n) {
        k_bdryeff_energy<<<dim3(nb256), dim3(256), 0, s>>>(ba);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    return PDM_OK;
}

}  // namespace pdm
