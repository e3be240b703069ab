// pdm_noalloc.hip -- the family force sums with no device allocation in the call.
//
// pdm_compute_forces copies a host field in and out through buffers it allocates.  Callers
// that cannot allocate (code under HIP graph capture, serving loops) pass device buffers in
// reference order to pdm_compute_forces_noalloc instead; the internal-order copies the tile
// gathers read and write live in one file-scope device array per GPU (`scratch`), whose
// address each device's first pdm_init looks up once.  Every call on a GPU shares that GPU's
// array, so calls that overlap on one GPU race; a call overwrites all of the scratch it uses
// before reading it, so nothing needs clearing between calls.
#include <mutex>

#include "pdm_internal.h"

namespace pdm {

// Internal-order copies of one pdm_compute_forces_noalloc call: u and F (d4) and the energy
// density (double) per node, up to kNoAllocMaxNodes nodes: 36 MiB on every GPU the library
// runs on.
constexpr size_t kNoAllocBytesPerNode = 2 * sizeof(d4) + sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytesPerNode * (size_t)kNoAllocMaxNodes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// u_int[a] = u_ref[perm[a]] (reference-order AoS doubles into internal-order d4).
__global__ __launch_bounds__(256) void k_to_internal(int64_t n, const int32_t* __restrict__ perm,
                                                     const double* __restrict__ u, d4* __restrict__ out)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int64_t i = perm[a];
    out[a] = d4{u[3 * i], u[3 * i + 1], u[3 * i + 2], 0.0};
}

// F_ref[perm[a]] = F_int[a], el_ref[perm[a]] = el_int[a] (el_ref may be null).
__global__ __launch_bounds__(256) void k_from_internal(int64_t n, const int32_t* __restrict__ perm,
                                                       const d4* __restrict__ F, const double* __restrict__ el,
                                                       double* __restrict__ Fo, double* __restrict__ elo)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int64_t i = perm[a];
    const d4 f = F[a];
    Fo[3 * i] = f.x;
    Fo[3 * i + 1] = f.y;
    Fo[3 * i + 2] = f.z;
    if (elo) elo[i] = el[a];
}

}  // namespace

hipError_t noalloc_scratch_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return hipSuccess;
    void* p = nullptr;
    // the caller has made dev current; the symbol's host-side name is not a device address
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
    return hipSuccess;
}

hipError_t noalloc_forces(const pdm_handle* h, const TileView& tv, const double* u, double* F, double* elast)
{
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (h->dev >= 0 && h->dev < kMaxDevices) base = g_scratch_ptr[h->dev];
    }
    if (!base) return hipErrorNotInitialized;
    const int64_t n = h->n;
    char* p = static_cast<char*>(base);
    d4* ui = reinterpret_cast<d4*>(p);
    d4* Fi = reinterpret_cast<d4*>(p + sizeof(d4) * (size_t)n);
    double* ei = reinterpret_cast<double*>(p + 2 * sizeof(d4) * (size_t)n);
    const unsigned g = (unsigned)((n + 255) / 256);
    k_to_internal<<<dim3(g), dim3(256), 0, h->stream>>>(n, h->permd, u, ui);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_forces(h->opt.precision, tv, ui, Fi, ei, h->stream);
    if (e != hipSuccess) return e;
    k_from_internal<<<dim3(g), dim3(256), 0, h->stream>>>(n, h->permd, Fi, ei, F, elast);
    return hipGetLastError();
}

}  // namespace pdm
