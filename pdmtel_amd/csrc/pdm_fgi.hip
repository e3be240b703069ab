// pdm_fgi.hip -- set-up of fixed Gaussian integration (fgi_on 1): Gauss points and the
// moving-least-squares (MLS) coefficients that interpolate their displacements.
//
// The reference builds, per node, n^3 Gauss points at X + (dx/2)(a_i, a_j, a_k) with weights
// w_i w_j w_k (SetGaussianPoint, pdDataManager2.cpp:72-104; rule of SetGaussianPointsValue,
// pdDataManager3.cpp:1706-1799).  For n > 1 each point's displacement is an MLS interpolation
// of the node's contributing nodes (InitIntegrationPoint, pdDataManager2.cpp:106-256): the
// node itself and its family within the support size rw (within 1.5 rw when that leaves
// fewer than 10), quadratic basis (GetBasis, :1547-1562), quartic spline weight of the
// distance from the point over rw (GetWeight, :1564-1577), phi_k = w_k p_k^T A^-1 p(x_g) with
// the moment matrix A = sum_k w_k p_k p_k^T.
//
// Here every table is built on the GPU in internal (tile) order.  The basis is evaluated
// relative to the node and in units of dx, where the reference uses global coordinates: the
// quadratic basis spans the same space after a shift and a scaling, so the shape functions
// are the same, while the moment matrix's condition number drops from about 1e12 to about 20.
// A is symmetric positive definite and is solved by Cholesky in registers.
#include <cmath>
#include <vector>

#include "pdm_geom.h"
#include "pdm_internal.h"

namespace pdm {
namespace {

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            h->err = std::string("HIP error ") + hipGetErrorString(e_) + " at " #x;  \
            return PDM_E_HIP;                                                       \
        }                                                                           \
    } while (0)

constexpr int kBasis = 10;  // quadratic basis in 3-D (BASISDIM, pdDataManager.h:37)

// Gauss-Legendre abscissae and weights of order n in the reference's point order (the zero
// first for odd n, then +a, -a pairs from the centre out; n = 8 lists the four positive
// points, then the four negative ones).  n = 2..5 in closed form, n = 6 and 8 from the
// standard tables.  false: the reference has no rule of that order.
bool gauss_rule(int n, double* a, double* w)
{
    switch (n) {
    case 1:
        a[0] = 0.0; w[0] = 2.0;
        return true;
    case 2:
        a[0] = 1.0 / std::sqrt(3.0); a[1] = -a[0];
        w[0] = w[1] = 1.0;
        return true;
    case 3:
        a[0] = 0.0; a[1] = std::sqrt(3.0 / 5.0); a[2] = -a[1];
        w[0] = 8.0 / 9.0; w[1] = w[2] = 5.0 / 9.0;
        return true;
    case 4: {
        const double r = 2.0 * std::sqrt(6.0 / 5.0);
        a[0] = std::sqrt((3.0 - r) / 7.0); a[1] = -a[0];
        a[2] = std::sqrt((3.0 + r) / 7.0); a[3] = -a[2];
        w[0] = w[1] = (18.0 + std::sqrt(30.0)) / 36.0;
        w[2] = w[3] = (18.0 - std::sqrt(30.0)) / 36.0;
        return true;
    }
    case 5: {
        const double r = 2.0 * std::sqrt(10.0 / 7.0);
        a[0] = 0.0;
        a[1] = std::sqrt(5.0 - r) / 3.0; a[2] = -a[1];
        a[3] = std::sqrt(5.0 + r) / 3.0; a[4] = -a[3];
        w[0] = 128.0 / 225.0;
        w[1] = w[2] = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        w[3] = w[4] = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0;
        return true;
    }
    case 6: {
        static const double x[3] = {0.2386191860831969086305017, 0.6612093864662645136613996,
                                    0.9324695142031520278123016};
        static const double c[3] = {0.4679139345726910473898703, 0.3607615730481386075698335,
                                    0.1713244923791703450402961};
        for (int k = 0; k < 3; ++k) {
            a[2 * k] = x[k]; a[2 * k + 1] = -x[k];
            w[2 * k] = w[2 * k + 1] = c[k];
        }
        return true;
    }
    case 8: {
        static const double x[4] = {0.1834346424956498049394761, 0.5255324099163289858177390,
                                    0.7966664774136267395915539, 0.9602898564975362316835609};
        static const double c[4] = {0.3626837833783619829651504, 0.3137066458778872873379622,
                                    0.2223810344533744705443560, 0.1012285362903762591525314};
        for (int k = 0; k < 4; ++k) {
            a[k] = x[k]; a[k + 4] = -x[k];
            w[k] = w[k + 4] = c[k];
        }
        return true;
    }
    default:
        return false;
    }
}

__device__ __forceinline__ void basis(double x, double y, double z, double* p)
{
    p[0] = 1.0; p[1] = x; p[2] = y; p[3] = z;
    p[4] = x * y; p[5] = x * z; p[6] = y * z;
    p[7] = x * x; p[8] = y * y; p[9] = z * z;
}

// Quartic spline weight of the distance from point xg to node xk over rw (GetWeight).
__device__ __forceinline__ double weight(const d4& xk, const double* xg, double rw)
{
    const double a = xg[0] - xk.x, b = xg[1] - xk.y, c = xg[2] - xk.z;
    const double d = sqrt(a * a + b * b + c * c) / rw;
    if (d > 1.0) return 0.0;
    return 1 - 6.0 * d * d + 8.0 * (d * d * d) - 3.0 * (d * d * d * d);
}

struct MlsArgs {
    const d4* Xd;        // reference order
    const int64_t* rp;   // reference-order families
    const int32_t* col;
    const int32_t* iperm;// reference -> internal id
    const d4* Xi;        // internal order
    int64_t n;
    double rw;
    int64_t* cnt;        // per internal node: contributing nodes
    uint8_t* wide;       // per reference node: 1 if the list uses 1.5 rw
    const int64_t* cb_ptr;
    int32_t* cb;
    double* phi;
    const double* tab;
    int ng;
    double dx;
    int* err;
};

// Contributing-node counts: the node and its family within rw, or within 1.5 rw when that
// gives fewer than kBasis nodes (pdDataManager2.cpp:118-160).  One thread per node.
__global__ void k_mls_count(MlsArgs m)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    const d4 xi = m.Xd[i];
    int64_t c = 1, cw = 1;
    for (int64_t e = m.rp[i]; e < m.rp[i + 1]; ++e) {
        const d4 xj = m.Xd[m.col[e]];
        const double a = xj.x - xi.x, b = xj.y - xi.y, z = xj.z - xi.z;
        const double d = sqrt(a * a + b * b + z * z);
        c += (d <= m.rw) ? 1 : 0;
        cw += (d <= m.rw * 1.5) ? 1 : 0;
    }
    const bool wide = c < kBasis;
    m.wide[i] = wide ? 1 : 0;
    m.cnt[m.iperm[i]] = wide ? cw : c;
}

// The contributing lists in internal ids, in the reference's order (the node, then its
// family in ascending reference id).
__global__ void k_mls_fill(MlsArgs m)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    const d4 xi = m.Xd[i];
    const double rad = m.wide[i] ? m.rw * 1.5 : m.rw;
    int64_t o = m.cb_ptr[m.iperm[i]];
    m.cb[o++] = m.iperm[i];
    for (int64_t e = m.rp[i]; e < m.rp[i + 1]; ++e) {
        const d4 xj = m.Xd[m.col[e]];
        const double a = xj.x - xi.x, b = xj.y - xi.y, z = xj.z - xi.z;
        if (sqrt(a * a + b * b + z * z) <= rad) m.cb[o++] = m.iperm[m.col[e]];
    }
}

__device__ __forceinline__ constexpr int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// MLS coefficients at the point xg of internal node a's contributing nodes: coefficient k of
// the list at out[k * stride].  The basis is taken relative to node a in units of dx.  false if
// the moment matrix is not positive definite.
__device__ bool mls_point(const MlsArgs& m, int64_t a, const double* xg, double* out, int64_t stride)
{
    const d4 xa = m.Xi[a];
    const int64_t k0 = m.cb_ptr[a], k1 = m.cb_ptr[a + 1];
    double A[kBasis * (kBasis + 1) / 2];
#pragma unroll
    for (int r = 0; r < kBasis * (kBasis + 1) / 2; ++r) A[r] = 0.0;
    for (int64_t k = k0; k < k1; ++k) {
        const d4 xk = m.Xi[m.cb[k]];
        double p[kBasis];
        basis((xk.x - xa.x) / m.dx, (xk.y - xa.y) / m.dx, (xk.z - xa.z) / m.dx, p);
        const double w = weight(xk, xg, m.rw);
#pragma unroll
        for (int r = 0; r < kBasis; ++r)
#pragma unroll
            for (int c = 0; c <= r; ++c) A[tri(r, c)] += w * p[r] * p[c];
    }
    // Cholesky A = L L^T in place
    bool ok = true;
#pragma unroll
    for (int j = 0; j < kBasis; ++j) {
        double d = A[tri(j, j)];
#pragma unroll
        for (int c = 0; c < j; ++c) d -= A[tri(j, c)] * A[tri(j, c)];
        ok = ok && d > 0.0;
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        A[tri(j, j)] = ljj;
#pragma unroll
        for (int r = j + 1; r < kBasis; ++r) {
            double v = A[tri(r, j)];
#pragma unroll
            for (int c = 0; c < j; ++c) v -= A[tri(r, c)] * A[tri(j, c)];
            A[tri(r, j)] = v / ljj;
        }
    }
    if (!ok) return false;
    double s[kBasis];
    basis((xg[0] - xa.x) / m.dx, (xg[1] - xa.y) / m.dx, (xg[2] - xa.z) / m.dx, s);
#pragma unroll
    for (int r = 0; r < kBasis; ++r) {
        double v = s[r];
#pragma unroll
        for (int c = 0; c < r; ++c) v -= A[tri(r, c)] * s[c];
        s[r] = v / A[tri(r, r)];
    }
#pragma unroll
    for (int r = kBasis - 1; r >= 0; --r) {
        double v = s[r];
#pragma unroll
        for (int c = r + 1; c < kBasis; ++c) v -= A[tri(c, r)] * s[c];
        s[r] = v / A[tri(r, r)];
    }
    for (int64_t k = k0; k < k1; ++k) {
        const d4 xk = m.Xi[m.cb[k]];
        double p[kBasis];
        basis((xk.x - xa.x) / m.dx, (xk.y - xa.y) / m.dx, (xk.z - xa.z) / m.dx, p);
        double v = 0.0;
#pragma unroll
        for (int r = 0; r < kBasis; ++r) v += p[r] * s[r];
        out[(k - k0) * stride] = weight(xk, xg, m.rw) * v;
    }
    return true;
}

// MLS coefficients of Gauss point g of internal node a (one thread per point).
__global__ __launch_bounds__(256) void k_mls_coef(MlsArgs m)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.n * m.ng) return;
    const int64_t a = q / m.ng;
    const int g = (int)(q - a * m.ng);
    const d4 xa = m.Xi[a];
    const double* t = m.tab + kFgiTab * g;
    const double xg[3] = {t[0] + xa.x, t[1] + xa.y, t[2] + xa.z};
    if (!mls_point(m, a, xg, m.phi + m.cb_ptr[a] * m.ng + g, m.ng)) *m.err = 1;
}

// ------------------------------------------------------------ adaptive integration (ai_on)
struct AiBuild {
    TileView tv;
    const int32_t* refid;    // internal -> reference id
    const int64_t* cb_ptr;   // MLS contributing nodes (mls) or null
    double horizon, dx;
    int nseg;
    int64_t* cnt;            // count pass: points, coefficients and entries with points, per node
    int64_t* cphi;
    int64_t* ent;
    const int64_t* ai_ptr;   // fill pass
    const int64_t* phi_base;
    d4* xw;
    uint32_t* eo;
    int64_t* po;             // the point's first coefficient; coefficient k at po + k * ps
    int32_t* ps;             // points of the point's entry (the coefficient stride)
    int32_t* pi;             // the point's nodes i and j
    int32_t* pj;
};

// The AI points of one entry of internal row i: node j's cell relative to node i, with the bond
// distances InitBond stores (from the lower reference id to the higher one's cell).
template <class EMIT>
__device__ __forceinline__ int ai_entry_points(const AiBuild& b, int64_t i, int32_t j, const d4& xi, const d4& xj,
                                               EMIT emit)
{
    double dmin, dmax;
    if (b.refid[i] < b.refid[j]) g_dist_to_cube_mm(xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, b.dx, &dmin, &dmax);
    else g_dist_to_cube_mm(xj.x, xj.y, xj.z, xi.x, xi.y, xi.z, b.dx, &dmin, &dmax);
    const double c[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
    return g_ai_cell_points(c, dmin, dmax, b.horizon, b.dx, b.nseg, emit);
}

// Count (FILL = false) or write (FILL = true) the points of every internal row, in row order,
// then x, y, z order within an entry.  One workgroup per tile, one wavefront per node, one lane
// per entry; cut entries and row padding (coefficient 0) have none.
template <bool FILL>
__global__ __launch_bounds__(256) void k_ai_pts(AiBuild b)
{
    const TileView& tv = b.tv;
    const int64_t k = blockIdx.x;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int lane = threadIdx.x & 63;
    for (int loc = threadIdx.x >> 6; loc < nt; loc += 4) {
        const int64_t i = a0 + loc;
        const d4 xi = tv.X[i];
        const int64_t eb = tv.rpi[i], ee = tv.rpi[i + 1];
        int64_t base = FILL ? b.ai_ptr[i] : 0, pbase = FILL ? b.phi_base[i] : 0;
        int64_t rc = 0, rp = 0, re = 0;
        for (int64_t e0 = eb; e0 < ee; e0 += 64) {
            const int64_t e = e0 + lane;
            const bool uncut = e < ee && tv.coef64[e] != 0.0;
            int32_t j = 0;
            d4 xj = xi;
            int c = 0;
            int64_t m = 0;
            if (uncut) {
                j = tv.halo[h0 + tv.slot[e]];
                xj = tv.X[j];
                c = ai_entry_points(b, i, j, xi, xj, [](int, double, double, double, double) {});
                m = b.cb_ptr ? b.cb_ptr[j + 1] - b.cb_ptr[j] : 0;
            }
            // inclusive prefix of the lanes' point and coefficient counts
            int64_t ic = c, ip = (int64_t)c * m;
            for (int off = 1; off < 64; off <<= 1) {
                const int64_t v = __shfl_up(ic, off, 64), w = __shfl_up(ip, off, 64);
                if (lane >= off) {
                    ic += v;
                    ip += w;
                }
            }
            const int64_t tc = __shfl(ic, 63, 64), tp = __shfl(ip, 63, 64);
            if (FILL && c > 0) {
                const int64_t o = base + (ic - c);
                const int64_t po = pbase + (ip - (int64_t)c * m);
                const uint32_t off = (uint32_t)(e - eb);
                ai_entry_points(b, i, j, xi, xj, [&](int q, double x, double y, double z, double w) {
                    b.xw[o + q] = d4{x, y, z, w};
                    b.eo[o + q] = off;
                    // coefficients of an entry: [contributing node][point], so that consecutive
                    // points read consecutive words
                    b.po[o + q] = po + q;
                    b.ps[o + q] = c;
                    b.pi[o + q] = (int32_t)i;
                    b.pj[o + q] = j;
                });
            }
            base += tc;
            pbase += tp;
            rc += tc;
            rp += tp;
            re += __popcll(__ballot(c > 0));
        }
        if (!FILL && lane == 0) {
            b.cnt[i] = rc;
            b.cphi[i] = rp;
            b.ent[i] = re;
        }
    }
}

// MLS coefficients of every AI point over its cell owner's (node j's) contributing nodes, at
// the point's global position X_i + xi (InitMLSCoeff, pdIntgManager2.cpp:1879-1930).
__global__ __launch_bounds__(256) void k_ai_mls(MlsArgs m, const d4* __restrict__ xw, const int64_t* __restrict__ po,
                                                const int32_t* __restrict__ ps, const int32_t* __restrict__ pi,
                                                const int32_t* __restrict__ pj, double* __restrict__ phi, int64_t np)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const d4 x = xw[q];
    const d4 xi = m.Xi[pi[q]];
    const double xg[3] = {x.x + xi.x, x.y + xi.y, x.z + xi.z};
    if (!mls_point(m, pj[q], xg, phi + po[q], ps[q])) *m.err = 1;
}

// Bytes above which the MLS coefficient table is refused.
constexpr double kMaxPhiBytes = 96.0 * 1024 * 1024 * 1024;
constexpr double kMaxPairBytes = 64.0 * 1024 * 1024 * 1024;

}  // namespace

// Device temporaries of the MLS set-up; they live in the caller's frame, registered with the
// caller's DevTemps, because the coefficient kernels read them after build_contrib returns.
struct ContribTemps {
    int32_t* ip = nullptr;
    int64_t* cnt = nullptr;
    uint8_t* wide = nullptr;
    int* err = nullptr;
};

// MLS contributing nodes of every node (h->cb_ptr, h->cb, h->ncb) and the argument block of
// the coefficient kernels; the temporaries are released with `temps`.
int build_contrib(pdm_handle* h, MlsArgs& m, DevTemps& temps, ContribTemps& ct)
{
    const Deck& d = h->deck;
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    std::vector<int32_t> iperm(n);
    for (int64_t a2 = 0; a2 < n; ++a2) iperm[h->perm[a2]] = (int32_t)a2;
    int32_t*& ip = ct.ip;
    int64_t*& cnt = ct.cnt;
    uint8_t*& wide = ct.wide;
    int*& err = ct.err;
    temps.add(&ct.ip);
    temps.add(&ct.cnt);
    temps.add(&ct.wide);
    temps.add(&ct.err);
    HIPCHK(hipMalloc(&ip, sizeof(int32_t) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMalloc(&cnt, sizeof(int64_t) * std::max<int64_t>(n, 1)));
    HIPCHK(hipMalloc(&wide, std::max<int64_t>(n, 1)));
    HIPCHK(hipMalloc(&err, sizeof(int)));
    HIPCHK(hipMemcpy(ip, iperm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(int), s));
    HIPCHK(hipMalloc(&h->cb_ptr, sizeof(int64_t) * (n + 1)));
    m.Xd = h->Xd;
    m.rp = h->rp;
    m.col = h->col;
    m.iperm = ip;
    m.Xi = h->Xi;
    m.n = n;
    m.rw = d.rw;
    m.cnt = cnt;
    m.wide = wide;
    m.dx = d.dx;
    m.err = err;
    const unsigned nb = (unsigned)((n + 255) / 256);
    k_mls_count<<<dim3(nb), dim3(256), 0, s>>>(m);
    HIPCHK(hipGetLastError());
    int rc = scan_exclusive(h, cnt, h->cb_ptr, n);
    if (rc) return rc;
    int64_t total = 0;
    HIPCHK(hipMemcpy(&total, h->cb_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost));
    h->ncb = total;
    HIPCHK(hipMalloc(&h->cb, sizeof(int32_t) * std::max<int64_t>(total, 1)));
    m.cb_ptr = h->cb_ptr;
    m.cb = h->cb;
    k_mls_fill<<<dim3(nb), dim3(256), 0, s>>>(m);
    HIPCHK(hipGetLastError());
    return PDM_OK;
}

void free_fgi(pdm_handle* h)
{
    for (void* p : {(void*)h->fgi_tab, (void*)h->gauss, (void*)h->fgi_trial, (void*)h->fgi_u0, (void*)h->cb_ptr,
                    (void*)h->cb, (void*)h->phi, (void*)h->fgi_pair_ptr, (void*)h->fgi_pairs})
        if (p) (void)hipFree(p);
    h->fgi_tab = nullptr;
    h->gauss = nullptr;
    h->fgi_trial = nullptr;
    h->fgi_u0 = nullptr;
    h->cb_ptr = nullptr;
    h->cb = nullptr;
    h->phi = nullptr;
    h->fgi_pair_ptr = nullptr;
    h->fgi_pairs = nullptr;
    h->fgi_npairs = 0;
}

// Gauss points, their table and (n > 1) the MLS lists and coefficients, after the tile build.
int build_fgi(pdm_handle* h)
{
    const Deck& d = h->deck;
    const int n1 = d.intg_point;
    double a[8], w[8];
    if (!gauss_rule(n1, a, w)) {
        h->err = "intg_point must be 1..6 or 8 (the reference's Gauss rules)";
        return PDM_E_PARSE;
    }
    const int ng = n1 * n1 * n1;
    h->fgi_n = n1;
    h->fgi_ng = ng;
    // per point: offsets (0.5 dx a), force weight, beta cell size, energy factor -- in the
    // reference's operation order (pdDataManager2.cpp:90-92, 1311; pdDataManager.cpp:1942-1954)
    std::vector<double> tab((size_t)kFgiTab * ng, 0.0);
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j)
            for (int k = 0; k < n1; ++k) {
                double* t = &tab[(size_t)kFgiTab * ((i * n1 + j) * n1 + k)];
                t[0] = 0.5 * d.dx * a[i];
                t[1] = 0.5 * d.dx * a[j];
                t[2] = 0.5 * d.dx * a[k];
                const double eff = std::pow(w[i] * w[j] * w[k], 1.0 / 3.0);
                t[3] = (std::pow(d.dx, 3) / 8.0) * w[i] * w[j] * w[k];
                t[4] = eff * d.dx / 2.0;
                t[5] = std::pow(0.5 * eff, 3);
            }
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    HIPCHK(hipMalloc(&h->fgi_tab, sizeof(double) * tab.size()));
    HIPCHK(hipMemcpy(h->fgi_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    const size_t gbytes = sizeof(d4) * (size_t)std::max<int64_t>(n, 1) * ng;
    HIPCHK(hipMalloc(&h->gauss, gbytes));
    HIPCHK(hipMemsetAsync(h->gauss, 0, gbytes, s));  // zero initial displacement (pdDataManager2.cpp:95-97)
    // fgi_trial: two buffers of N (the stages alternate between them), fgi_u0: N
    for (d4** p : {&h->fgi_trial, &h->fgi_u0}) {
        const size_t bytes = sizeof(d4) * std::max<int64_t>(n, 1) * (p == &h->fgi_trial ? 2 : 1);
        HIPCHK(hipMalloc(p, bytes));
        HIPCHK(hipMemsetAsync(*p, 0, bytes, s));
    }
    {
        // live (entry, Gauss point) pairs: the gathers skip the pairs whose beta is 0
        HIPCHK(hipMalloc(&h->fgi_pair_ptr, sizeof(int64_t) * (n + 1)));
        int64_t *cnt = nullptr, *ent = nullptr;
        HIPCHK(hipMalloc(&cnt, sizeof(int64_t) * std::max<int64_t>(n, 1)));
        HIPCHK(hipMalloc(&ent, sizeof(int64_t) * std::max<int64_t>(n, 1)));
        const TileView tv = tile_view(h);
        FgiArgs f = fgi_args(h);
        hipError_t e = launch_fgi_pairs(tv, f, false, cnt, ent, nullptr, s);
        int rc = (e == hipSuccess) ? scan_exclusive(h, cnt, h->fgi_pair_ptr, n) : PDM_E_HIP;
        std::vector<int64_t> enth(n);
        if (rc == PDM_OK && n) {
            e = hipMemcpy(enth.data(), ent, sizeof(int64_t) * n, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(&h->fgi_npairs, h->fgi_pair_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost);
        }
        (void)hipFree(cnt);
        (void)hipFree(ent);
        if (e != hipSuccess) {
            h->err = std::string("HIP error ") + hipGetErrorString(e) + " in the Gauss-point pair lists";
            return PDM_E_HIP;
        }
        if (rc) return rc;
        if ((double)h->fgi_npairs * sizeof(uint2) > kMaxPairBytes) {
            h->err = "Gauss-point pair lists exceed 64 GiB (lower intg_point)";
            return PDM_E_LIMIT;
        }
        h->Elive = 0;
        for (int64_t a2 = 0; a2 < n; ++a2) h->Elive += enth[a2];
        HIPCHK(hipMalloc(&h->fgi_pairs, sizeof(uint2) * std::max<int64_t>(h->fgi_npairs, 1)));
        f = fgi_args(h);
        HIPCHK(launch_fgi_pairs(tv, f, true, nullptr, nullptr, h->fgi_pairs, s));
    }
    if (n1 == 1 || n == 0) {
        HIPCHK(hipStreamSynchronize(s));
        return PDM_OK;
    }
    MlsArgs m{};
    ContribTemps ct;
    DevTemps temps(s);
    int rc = build_contrib(h, m, temps, ct);
    if (rc) return rc;
    const int64_t total = h->ncb;
    if ((double)total * ng * sizeof(double) > kMaxPhiBytes) {
        h->err = "MLS coefficient table exceeds 96 GiB (lower intg_point or support_domain_size)";
        return PDM_E_LIMIT;
    }
    HIPCHK(hipMalloc(&h->phi, sizeof(double) * std::max<int64_t>(total, 1) * ng));
    m.phi = h->phi;
    m.tab = h->fgi_tab;
    m.ng = ng;
    const int64_t q = n * ng;
    k_mls_coef<<<dim3((unsigned)((q + 255) / 256)), dim3(256), 0, s>>>(m);
    HIPCHK(hipGetLastError());
    int herr = 0;
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(&herr, m.err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        h->err = "MLS moment matrix is singular at a Gauss point (too few contributing nodes)";
        return PDM_E_MODEL;
    }
    return PDM_OK;
}

void free_ai(pdm_handle* h)
{
    for (void* p : {(void*)h->ai_ptr, (void*)h->ai_xw, (void*)h->ai_eo, (void*)h->ai_po, (void*)h->ai_phi,
                    (void*)h->ai_j, (void*)h->ai_i, (void*)h->ai_ps, (void*)h->ai_res})
        if (p) (void)hipFree(p);
    h->ai_j = h->ai_i = h->ai_ps = nullptr;
    h->ai_res = nullptr;
    h->ai_ptr = nullptr;
    h->ai_xw = nullptr;
    h->ai_eo = nullptr;
    h->ai_po = nullptr;
    h->ai_phi = nullptr;
    h->ai_np = h->ai_nphi = 0;
}

// Adaptive integration (ai_on 1) after the tile build: the trapezoidal points of every entry
// (InitIntegrationPoint's AI branch, pdDataManager2.cpp:257-297, places them with a unit
// integrand) and, with MLS, their coefficients over node j's contributing nodes.  Without MLS
// the integrand prescribes each point's relative displacement from node j's DISPLGRAD
// condition and exits when node j has none (pdIntgManager2.cpp:187-211): such a deck is refused
// here unless every node carries one.
int build_ai(pdm_handle* h)
{
    const Deck& d = h->deck;
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    h->ai = d.mls_on ? 1 : 2;
    h->ai_nseg = 1 << (d.intg_point - 1);
    if (!d.mls_on)
        for (int64_t i = 0; i < n; ++i) {
            const int b = h->bc[i];
            if (b < 0 || d.bcs[b].type != 3) {
                h->err = "adaptive integration without MLS (ai_on 1 0) needs a DISPLGRAD condition on every node";
                return PDM_E_MODEL;
            }
        }
    // fgi_trial: two buffers of N (the stages alternate between them), fgi_u0: N
    for (d4** p : {&h->fgi_trial, &h->fgi_u0}) {
        const size_t bytes = sizeof(d4) * std::max<int64_t>(n, 1) * (p == &h->fgi_trial ? 2 : 1);
        HIPCHK(hipMalloc(p, bytes));
        HIPCHK(hipMemsetAsync(*p, 0, bytes, s));
    }
    HIPCHK(hipMalloc(&h->ai_ptr, sizeof(int64_t) * (n + 1)));
    if (n == 0) {
        HIPCHK(hipMemsetAsync(h->ai_ptr, 0, sizeof(int64_t), s));
        HIPCHK(hipStreamSynchronize(s));
        return PDM_OK;
    }
    MlsArgs m{};
    ContribTemps ct;
    DevTemps temps(s);
    if (d.mls_on) {
        int rc = build_contrib(h, m, temps, ct);
        if (rc) return rc;
    }
    AiBuild b{};
    b.tv = tile_view(h);
    b.refid = h->permd;
    b.cb_ptr = d.mls_on ? h->cb_ptr : nullptr;
    b.horizon = d.horizon;
    b.dx = d.dx;
    b.nseg = h->ai_nseg;
    int64_t *cnt = nullptr, *cphi = nullptr, *ent = nullptr, *phib = nullptr;
    temps.add(&cnt);
    temps.add(&cphi);
    temps.add(&ent);
    temps.add(&phib);
    HIPCHK(hipMalloc(&cnt, sizeof(int64_t) * n));
    HIPCHK(hipMalloc(&cphi, sizeof(int64_t) * n));
    HIPCHK(hipMalloc(&ent, sizeof(int64_t) * n));
    HIPCHK(hipMalloc(&phib, sizeof(int64_t) * (n + 1)));
    b.cnt = cnt;
    b.cphi = cphi;
    b.ent = ent;
    k_ai_pts<false><<<dim3((unsigned)h->ntl), dim3(256), 0, s>>>(b);
    HIPCHK(hipGetLastError());
    int rc = scan_exclusive(h, cnt, h->ai_ptr, n);
    if (rc) return rc;
    rc = scan_exclusive(h, cphi, phib, n);
    if (rc) return rc;
    std::vector<int64_t> enth(n);
    HIPCHK(hipMemcpy(enth.data(), ent, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&h->ai_np, h->ai_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&h->ai_nphi, phib + n, sizeof(int64_t), hipMemcpyDeviceToHost));
    h->Elive = 0;
    for (int64_t a2 = 0; a2 < n; ++a2) h->Elive += enth[a2];
    const double bytes = (double)h->ai_np * (sizeof(d4) + sizeof(uint32_t) + sizeof(int64_t) + 3 * sizeof(int32_t)) +
                         (double)h->ai_nphi * sizeof(double);
    if (bytes > kMaxPhiBytes) {
        h->err = "adaptive-integration point tables exceed 96 GiB (lower intg_point or support_domain_size)";
        return PDM_E_LIMIT;
    }
    const int64_t np1 = std::max<int64_t>(h->ai_np, 1);
    HIPCHK(hipMalloc(&h->ai_xw, sizeof(d4) * np1));
    HIPCHK(hipMalloc(&h->ai_eo, sizeof(uint32_t) * np1));
    HIPCHK(hipMalloc(&h->ai_po, sizeof(int64_t) * np1));
    HIPCHK(hipMalloc(&h->ai_ps, sizeof(int32_t) * np1));
    HIPCHK(hipMalloc(&h->ai_i, sizeof(int32_t) * np1));
    HIPCHK(hipMalloc(&h->ai_j, sizeof(int32_t) * np1));
    HIPCHK(hipMalloc(&h->ai_res, sizeof(d4) * n));
    HIPCHK(hipMemsetAsync(h->ai_res, 0, sizeof(d4) * n, s));
    b.ai_ptr = h->ai_ptr;
    b.phi_base = phib;
    b.xw = h->ai_xw;
    b.eo = h->ai_eo;
    b.po = h->ai_po;
    b.ps = h->ai_ps;
    b.pi = h->ai_i;
    b.pj = h->ai_j;
    k_ai_pts<true><<<dim3((unsigned)h->ntl), dim3(256), 0, s>>>(b);
    HIPCHK(hipGetLastError());
    if (d.mls_on) {
        HIPCHK(hipMalloc(&h->ai_phi, sizeof(double) * std::max<int64_t>(h->ai_nphi, 1)));
        if (h->ai_np)
            k_ai_mls<<<dim3((unsigned)((h->ai_np + 255) / 256)), dim3(256), 0, s>>>(m, h->ai_xw, h->ai_po, h->ai_ps, h->ai_i, h->ai_j,
                                                                                 h->ai_phi, h->ai_np);
        HIPCHK(hipGetLastError());
        int herr = 0;
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipMemcpy(&herr, m.err, sizeof(int), hipMemcpyDeviceToHost));
        if (herr) {
            h->err = "MLS moment matrix is singular at an integration point (too few contributing nodes)";
            return PDM_E_MODEL;
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    return PDM_OK;
}

FgiArgs fgi_args(const pdm_handle* h)
{
    FgiArgs f{};
    f.tab = h->fgi_tab;
    f.ng = h->fgi_ng;
    f.mode = 0;
    f.horizon = h->deck.horizon;
    f.gin = h->gauss;
    f.gout = h->gauss;
    f.cb_ptr = h->cb_ptr;
    f.cb = h->cb;
    f.phi = h->phi;
    f.X = h->Xi;
    f.bc = h->bci;
    f.bcs = h->bcs_d;
    f.steps = h->steps;
    f.n = h->n;
    f.trial = h->fgi_trial;
    f.u0 = h->fgi_u0;
    f.pair_ptr = h->fgi_pair_ptr;
    f.pairs = h->fgi_pairs;
    f.prec = h->opt.precision;
    f.ai = h->ai;
    f.ai_ptr = h->ai_ptr;
    f.ai_xw = h->ai_xw;
    f.ai_eo = h->ai_eo;
    f.ai_po = h->ai_po;
    f.ai_ps = h->ai_ps;
    f.ai_i = h->ai_i;
    f.ai_j = h->ai_j;
    f.ai_res = h->ai_res;
    f.ai_np = h->ai_np;
    f.ai_phi = h->ai_phi;
    f.ai_vol = h->deck.dx * h->deck.dx * h->deck.dx;
    return f;
}

}  // namespace pdm
