// pdm_cg.hip -- static equilibrium on the tangent of the bond operator.
//
// The reference assembles a dense (3N)^2 system from the linearised bond force
// and solves it by Gauss-Jordan (pdSolve.cpp:132-673, broken as shipped).  Here
// the same system is solved matrix-free:
//   K u = b,  (K u)_i = sum_j coef_ij xi (xi . (u_i - u_j)) / r^3
// with coef the force coefficient (bdryEff * dx^3 * beta * c), i.e. the tangent
// of the CCI force at eta = 0, so the static solution is the rest state of the
// dynamic path.  Prescribed rows (AssembleEquation, pdSolve.cpp:227-363): DISPL
// flagged components take value * factor(T), VELOC ones value * T, DISPLGRAD ones
// (G X) * factor(T), T = total_time; the rest are unknowns.  Load b: the body
// force counted once per unbroken bond, as the dynamic path does, plus the
// traction body force of the extension.
//
// Solver: block-Jacobi-preconditioned CG on the free components (mask m): each node's 3x3
// tangent block, restricted to its free components, is inverted once; all vector
// algebra in FP64, matvecs in the handle's precision (FP64, or FP32 bonds for
// PDM_FP32C/C2).  After every block of 25 iterations (captured in a hipGraph, CG
// scalars on the device) the residual is recomputed with the FP64 operator and
// replaces the recurrence residual, the search direction kept ("reliable update"
// residual replacement): the stopping test is on the FP64 residual, and FP32 matvecs
// converge to the FP64 tolerance without restarts.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

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

constexpr int kVB = 256;          // vector-kernel block
constexpr int kBlockIters = 25;   // CG iterations per captured graph

struct Scal {
    double rho, rho_new, pq, rr, rr0, beta;
};

__device__ __forceinline__ double dot3(const d4& a, const d4& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// z = m (B^-1 r): the node's inverse block, diagonal in ia, off-diagonal (12, 13, 23) in ib.
__device__ __forceinline__ d4 precond(const d4& r, const d4& ia, const d4& ib, const d4& m)
{
    return d4{m.x * (ia.x * r.x + ib.x * r.y + ib.y * r.z), m.y * (ib.x * r.x + ia.y * r.y + ib.z * r.z),
              m.z * (ib.y * r.x + ib.z * r.y + ia.z * r.z), 0.0};
}

__device__ __forceinline__ void block_sum2(double a, double b, double* out2)
{
    __shared__ double sa[kVB], sb[kVB];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int off = kVB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) {
            sa[threadIdx.x] += sa[threadIdx.x + off];
            sb[threadIdx.x] += sb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out2[2 * blockIdx.x] = sa[0];
        out2[2 * blockIdx.x + 1] = sb[0];
    }
}

// r = m (rhs - Ax); partial |r|^2.  (Ax may be null: r = m rhs.)
__global__ __launch_bounds__(kVB) void k_residual(int64_t n, const d4* __restrict__ rhs, const d4* __restrict__ Ax,
                                                  const d4* __restrict__ m, d4* __restrict__ r, double* part)
{
    int64_t i = (int64_t)blockIdx.x * kVB + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        d4 v = rhs[i];
        if (Ax) {
            d4 a = Ax[i];
            v.x -= a.x;
            v.y -= a.y;
            v.z -= a.z;
        }
        d4 mm = m[i];
        v = d4{v.x * mm.x, v.y * mm.y, v.z * mm.z, 0.0};
        r[i] = v;
        rr = dot3(v, v);
    }
    block_sum2(rr, 0.0, part);
}

// d = 0, z = m B^-1 r, p = z; partial r.z and |r|^2.
__global__ __launch_bounds__(kVB) void k_pcg_start(int64_t n, const d4* __restrict__ r, const d4* __restrict__ ia,
                                                   const d4* __restrict__ ib, const d4* __restrict__ m,
                                                   d4* __restrict__ z, d4* __restrict__ p, d4* __restrict__ d,
                                                   double* part)
{
    int64_t i = (int64_t)blockIdx.x * kVB + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        d4 ri = r[i];
        d4 zi = precond(ri, ia[i], ib[i], m[i]);
        z[i] = zi;
        p[i] = zi;
        d[i] = d4{0.0, 0.0, 0.0, 0.0};
        rz = dot3(ri, zi);
        rr = dot3(ri, ri);
    }
    block_sum2(rz, rr, part);
}

// alpha = rho / pq; d += alpha p; r -= alpha m q; z = m B^-1 r; partial r.z, |r|^2.
__global__ __launch_bounds__(kVB) void k_pcg_update(int64_t n, const Scal* __restrict__ sc, const d4* __restrict__ p,
                                                    const d4* __restrict__ q, const d4* __restrict__ ia,
                                                    const d4* __restrict__ ib, const d4* __restrict__ m,
                                                    d4* __restrict__ d, d4* __restrict__ r, d4* __restrict__ z,
                                                    double* part)
{
    int64_t i = (int64_t)blockIdx.x * kVB + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    const double alpha = (sc->pq != 0.0) ? sc->rho / sc->pq : 0.0;
    if (i < n) {
        d4 pi = p[i], qi = q[i], mm = m[i];
        d4 di = d[i], ri = r[i];
        di.x += alpha * pi.x;
        di.y += alpha * pi.y;
        di.z += alpha * pi.z;
        ri.x -= alpha * mm.x * qi.x;
        ri.y -= alpha * mm.y * qi.y;
        ri.z -= alpha * mm.z * qi.z;
        d4 zi = precond(ri, ia[i], ib[i], mm);
        d[i] = di;
        r[i] = ri;
        z[i] = zi;
        rz = dot3(ri, zi);
        rr = dot3(ri, ri);
    }
    block_sum2(rz, rr, part);
}

// z = m B^-1 r after a residual replacement; partial r.z, |r|^2 (p is kept).
__global__ __launch_bounds__(kVB) void k_pcg_refresh(int64_t n, const d4* __restrict__ r, const d4* __restrict__ ia,
                                                     const d4* __restrict__ ib, const d4* __restrict__ m,
                                                     d4* __restrict__ z, double* part)
{
    int64_t i = (int64_t)blockIdx.x * kVB + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        d4 ri = r[i];
        d4 zi = precond(ri, ia[i], ib[i], m[i]);
        z[i] = zi;
        rz = dot3(ri, zi);
        rr = dot3(ri, ri);
    }
    block_sum2(rz, rr, part);
}

// p = z + beta p (beta = rho_new / rho, set by the reduction that produced rho_new).
__global__ __launch_bounds__(kVB) void k_pcg_dir(int64_t n, const Scal* __restrict__ sc, const d4* __restrict__ z,
                                                 d4* __restrict__ p)
{
    int64_t i = (int64_t)blockIdx.x * kVB + threadIdx.x;
    const double beta = sc->beta;
    if (i < n) {
        d4 zi = z[i], pi = p[i];
        p[i] = d4{zi.x + beta * pi.x, zi.y + beta * pi.y, zi.z + beta * pi.z, 0.0};
    }
}

// x += d; out = x + uc (when out is set)
__global__ __launch_bounds__(kVB) void k_axpy_out(int64_t n, d4* __restrict__ x, const d4* __restrict__ d,
                                                  const d4* __restrict__ uc, d4* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * kVB + threadIdx.x;
    if (i >= n) return;
    d4 xi = x[i];
    if (d) {
        d4 di = d[i];
        xi = d4{xi.x + di.x, xi.y + di.y, xi.z + di.z, 0.0};
        x[i] = xi;
    }
    if (out) {
        d4 c = uc[i];
        out[i] = d4{xi.x + c.x, xi.y + c.y, xi.z + c.z, 0.0};
    }
}

// Reduce partial pairs and store into the scalar block.  mode 0: rr0 = sum a;
// 1: rho = sum a, rr = sum b; 2: pq = sum a; 3: rho_new = sum a, rr = sum b, beta =
// rho_new / rho for the direction update, and rho rolls to rho_new.
__global__ __launch_bounds__(kVB) void k_reduce_scal(const double* __restrict__ part, int64_t nb, Scal* sc, int mode)
{
    __shared__ double sa[kVB], sb[kVB];
    double a = 0.0, b = 0.0;
    for (int64_t q = threadIdx.x; q < nb; q += kVB) {
        a += part[2 * q];
        b += part[2 * q + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int off = kVB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) {
            sa[threadIdx.x] += sa[threadIdx.x + off];
            sb[threadIdx.x] += sb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == 0) sc->rr0 = sa[0];
        else if (mode == 1) { sc->rho = sa[0]; sc->rr = sb[0]; }
        else if (mode == 2) sc->pq = sa[0];
        else {
            sc->rho_new = sa[0];
            sc->rr = sb[0];
            sc->beta = (sc->rho != 0.0) ? sa[0] / sc->rho : 0.0;
            sc->rho = sa[0];
        }
    }
}

struct Work {
    d4 *m = nullptr, *uc = nullptr, *b = nullptr, *dg = nullptr, *og = nullptr, *rhs = nullptr, *x = nullptr,
       *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *d = nullptr, *zero = nullptr;
    double* part = nullptr;
    Scal* sc = nullptr;
    void release()
    {
        for (d4** v : {&m, &uc, &b, &dg, &og, &rhs, &x, &r, &z, &p, &q, &d, &zero})
            if (*v) (void)hipFree(*v);
        if (part) (void)hipFree(part);
        if (sc) (void)hipFree(sc);
    }
};

}  // namespace

int solve_static(pdm_handle* h, double rel_tol, int max_iterations, pdm_static_stats* stats)
{
    auto t0 = std::chrono::steady_clock::now();
    const Deck& d = h->deck;
    hipStream_t s = h->stream;
    const int64_t n = h->n;
    if (n == 0) return PDM_OK;
    if (!(rel_tol > 0.0)) rel_tol = 1e-10;
    if (max_iterations <= 0) max_iterations = 100000;
    // prescribed components, values and loads (internal order)
    const double T = d.total_time;
    std::vector<d4> mh(n), uch(n), bh(n);
    std::vector<int32_t> nbh(n);
    HIPCHK(hipMemcpy(nbh.data(), h->nb, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    for (int64_t a = 0; a < n; ++a) {
        const int64_t i = h->perm[a];
        const int m = h->mat[i];
        const int bc = h->bc[i];
        double mk[3] = {1.0, 1.0, 1.0}, uc[3] = {0.0, 0.0, 0.0}, bb[3];
        for (int c = 0; c < 3; ++c) bb[c] = (double)nbh[a] * (d.mats[m].g[c] * d.mats[m].rho);
        if (bc >= 0) {
            const BoundaryCondition& B = d.bcs[bc];
            const double fac = time_factor(B, T);
            const double* X = &h->X[3 * i];
            for (int c = 0; c < 3; ++c) {
                if (B.type == 4) bb[c] += h->trac_density[bc][c] * fac;
                if (B.f[c] != 1 || !is_bc_node(d, bc)) continue;
                if (B.type == 1) { mk[c] = 0.0; uc[c] = B.v[c] * fac; }
                else if (B.type == 2) { mk[c] = 0.0; uc[c] = B.v[c] * T; }
                else if (B.type == 3) { mk[c] = 0.0; uc[c] = (B.v[3 * c] * X[0] + B.v[3 * c + 1] * X[1] + B.v[3 * c + 2] * X[2]) * fac; }
            }
        }
        mh[a] = d4{mk[0], mk[1], mk[2], 0.0};
        uch[a] = d4{uc[0], uc[1], uc[2], 0.0};
        bh[a] = d4{bb[0], bb[1], bb[2], 0.0};
    }
    Work w;
    const size_t vb = sizeof(d4) * n;
    const int64_t nblk = (n + kVB - 1) / kVB;
    auto fail = [&](int code, const std::string& m) {
        h->err = m;
        w.release();
        return code;
    };
#define WCHK(x)                                                                              \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) return fail(PDM_E_HIP, std::string("HIP error ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)
    for (d4** v : {&w.m, &w.uc, &w.b, &w.dg, &w.og, &w.rhs, &w.x, &w.r, &w.z, &w.p, &w.q, &w.d, &w.zero})
        WCHK(hipMalloc(v, vb));
    const int64_t ntl = h->ntl;
    WCHK(hipMalloc(&w.part, sizeof(double) * 2 * std::max<int64_t>(nblk, std::max<int64_t>(ntl, 1))));
    WCHK(hipMalloc(&w.sc, sizeof(Scal)));
    WCHK(hipMemcpyAsync(w.m, mh.data(), vb, hipMemcpyHostToDevice, s));
    WCHK(hipMemcpyAsync(w.uc, uch.data(), vb, hipMemcpyHostToDevice, s));
    WCHK(hipMemcpyAsync(w.b, bh.data(), vb, hipMemcpyHostToDevice, s));
    WCHK(hipMemsetAsync(w.zero, 0, vb, s));
    WCHK(hipMemsetAsync(w.x, 0, vb, s));
    WCHK(hipMemsetAsync(w.sc, 0, sizeof(Scal), s));
    TileView tv;
    {
        const int rc = tangent_view(h, tv);  // the operator's per-entry weights coef / r^3
        if (rc) {
            w.release();
            return rc;
        }
    }
    const dim3 g((unsigned)nblk), bk(kVB);
    // Block-Jacobi preconditioner: each node's 3x3 tangent block sum_j w xi xi^T restricted to
    // its free components, inverted.  A free component with no bonds keeps diagonal 1 (its
    // row is empty); a block that is not safely invertible (all of a node's bonds on one line)
    // falls back to its diagonal.  w.dg / w.og then hold the inverse's diagonal and
    // off-diagonal (12, 13, 23) entries.
    WCHK(launch_tangent_diag(tv, w.zero, w.dg, s, w.og));
    {
        std::vector<d4> dgh(n), ogh(n);
        WCHK(hipMemcpyAsync(dgh.data(), w.dg, vb, hipMemcpyDeviceToHost, s));
        WCHK(hipMemcpyAsync(ogh.data(), w.og, vb, hipMemcpyDeviceToHost, s));
        WCHK(hipStreamSynchronize(s));
        for (int64_t a = 0; a < n; ++a) {
            const d4 mm = mh[a];
            const double fr[3] = {mm.x, mm.y, mm.z};
            double B[3][3] = {{dgh[a].x, ogh[a].x, ogh[a].y}, {ogh[a].x, dgh[a].y, ogh[a].z}, {ogh[a].y, ogh[a].z, dgh[a].z}};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    if (r != c && (fr[r] == 0.0 || fr[c] == 0.0)) B[r][c] = 0.0;
            for (int r = 0; r < 3; ++r)
                if (fr[r] == 0.0 || !(B[r][r] > 0.0)) {
                    for (int c = 0; c < 3; ++c) B[r][c] = B[c][r] = 0.0;
                    B[r][r] = 1.0;
                }
            const double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
            const double c01 = B[0][2] * B[2][1] - B[0][1] * B[2][2];
            const double c02 = B[0][1] * B[1][2] - B[0][2] * B[1][1];
            const double det = B[0][0] * c00 + B[1][0] * c01 + B[2][0] * c02;
            const double scale = B[0][0] * B[1][1] * B[2][2];
            if (det > 1e-8 * scale) {
                const double id = 1.0 / det;
                dgh[a] = d4{c00 * id, (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * id,
                            (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * id, 0.0};
                ogh[a] = d4{c01 * id, c02 * id, (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * id, 0.0};
            } else {
                dgh[a] = d4{1.0 / B[0][0], 1.0 / B[1][1], 1.0 / B[2][2], 0.0};
                ogh[a] = d4{0.0, 0.0, 0.0, 0.0};
            }
        }
        WCHK(hipMemcpyAsync(w.dg, dgh.data(), vb, hipMemcpyHostToDevice, s));
        WCHK(hipMemcpyAsync(w.og, ogh.data(), vb, hipMemcpyHostToDevice, s));
    }
    // rhs = m (b - K uc)
    WCHK(launch_tangent(PDM_FP64, tv, w.uc, w.q, s));
    k_residual<<<g, bk, 0, s>>>(n, w.b, w.q, w.m, w.rhs, w.part);
    WCHK(hipGetLastError());
    k_reduce_scal<<<1, kVB, 0, s>>>(w.part, nblk, w.sc, 0);
    Scal sh;
    WCHK(hipMemcpyAsync(&sh, w.sc, sizeof(Scal), hipMemcpyDeviceToHost, s));
    WCHK(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(sh.rr0);
    const bool mixed = h->opt.precision != PDM_FP64;
    const int inner_prec = h->opt.precision;  // PDM_FP32C: FP32 weights; PDM_FP32C2: FP32 + 16-bit correction
    int iters = 0, sweeps = 0;
    double rel = 0.0;
    hipGraphExec_t gexec = nullptr;
    // one captured block of kBlockIters PCG iterations
    auto enqueue_block = [&]() -> hipError_t {
        for (int k = 0; k < kBlockIters; ++k) {
            // q = K p, with each tile's share of p . q from the same kernel
            hipError_t e = launch_tangent(inner_prec, tv, w.p, w.q, s, w.part);
            if (e != hipSuccess) return e;
            k_reduce_scal<<<1, kVB, 0, s>>>(w.part, ntl, w.sc, 2);
            k_pcg_update<<<g, bk, 0, s>>>(n, w.sc, w.p, w.q, w.dg, w.og, w.m, w.d, w.r, w.z, w.part);
            k_reduce_scal<<<1, kVB, 0, s>>>(w.part, nblk, w.sc, 3);
            k_pcg_dir<<<g, bk, 0, s>>>(n, w.sc, w.z, w.p);
        }
        return hipGetLastError();
    };
    if (bnorm == 0.0) {
        rel = 0.0;
    } else {
        // One PCG on A d = rhs.  After every block of iterations the recurrence residual is
        // replaced by the residual rhs - A d evaluated with the FP64 operator, keeping the
        // search direction: with FP32 matvecs (PDM_FP32C) the solution still converges to
        // the FP64 tolerance, without the restarts of a refinement loop.
        k_residual<<<g, bk, 0, s>>>(n, w.rhs, nullptr, w.m, w.r, w.part);
        k_pcg_start<<<g, bk, 0, s>>>(n, w.r, w.dg, w.og, w.m, w.z, w.p, w.d, w.part);
        k_reduce_scal<<<1, kVB, 0, s>>>(w.part, nblk, w.sc, 1);
        WCHK(hipGetLastError());
        hipGraph_t gr;
        WCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t ce = enqueue_block();
        hipError_t ee = hipStreamEndCapture(s, &gr);
        WCHK(ce);
        WCHK(ee);
        WCHK(hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0));
        WCHK(hipGraphDestroy(gr));
        rel = 1.0;
        while (iters < max_iterations) {
            WCHK(hipGraphLaunch(gexec, s));
            iters += kBlockIters;
            // the FP64 residual of d replaces the recurrence one
            WCHK(launch_tangent(PDM_FP64, tv, w.d, w.q, s));
            k_residual<<<g, bk, 0, s>>>(n, w.rhs, w.q, w.m, w.r, w.part);
            k_reduce_scal<<<1, kVB, 0, s>>>(w.part, nblk, w.sc, 0);
            WCHK(hipMemcpyAsync(&sh, w.sc, sizeof(Scal), hipMemcpyDeviceToHost, s));
            WCHK(hipStreamSynchronize(s));
            rel = std::sqrt(sh.rr0) / bnorm;
            ++sweeps;
            if (h->opt.verbose && sweeps % 20 == 0)
                fprintf(stderr, "pdm_solve_static: %d iterations, relative residual %.3e\n", iters, rel);
            if (!std::isfinite(rel) || rel <= rel_tol) break;
            k_pcg_refresh<<<g, bk, 0, s>>>(n, w.r, w.dg, w.og, w.m, w.z, w.part);
            k_reduce_scal<<<1, kVB, 0, s>>>(w.part, nblk, w.sc, 1);
            WCHK(hipGetLastError());
        }
        k_axpy_out<<<g, bk, 0, s>>>(n, w.x, w.d, nullptr, nullptr);
        WCHK(hipGetLastError());
    }
    // U = x + uc, V = 0, and the next stage gathers U
    k_axpy_out<<<g, bk, 0, s>>>(n, w.x, nullptr, w.uc, h->U);
    WCHK(hipGetLastError());
    WCHK(hipMemcpyAsync(h->u[0], h->U, vb, hipMemcpyDeviceToDevice, s));
    WCHK(hipMemsetAsync(h->V, 0, vb, s));
    WCHK(hipStreamSynchronize(s));
    if (gexec) (void)hipGraphExecDestroy(gexec);
    w.release();
    if (stats) {
        stats->iterations = iters;
        stats->refinements = sweeps;
        stats->rel_residual = rel;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
#undef WCHK
    if (!(rel <= rel_tol)) {
        h->err = "static solve stopped at relative residual " + std::to_string(rel) + " after " +
                 std::to_string(iters) + " iterations";
        return PDM_E_CONVERGE;
    }
    return PDM_OK;
}

}  // namespace pdm
