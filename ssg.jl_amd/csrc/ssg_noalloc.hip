// Allocation-free tendency: compute_tendency! (src/nonlinear.jl:22-36) of
// device-resident b and φₛ for callers that cannot allocate device memory
// (code under stream capture, tight serving loops).  Every intermediate --
// twiddle factors, the surface velocities and the two spectral planes --
// lives in one file-scope device array; the call enqueues its kernels on the
// caller's stream and returns without synchronising.  Declaration:
// include/ssg.h (ssg_tendency_noalloc).
#include "ssg_internal.hpp"

namespace ssg_impl {
namespace {

// Scratch of the allocation-free path: 40 MiB per device, held until the
// process ends.  An FP64 grid of nx*ny points needs twiddles (nx + ny), the
// velocity pair (nx*ny) and two planes ((nx/2+1)*ny), all complex FP64:
// 1024 x 1024 needs 33.6 MB.
constexpr size_t kNoallocScratchBytes = size_t(40) << 20;
constexpr size_t kAlign = 256;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
thread_local char g_noalloc_err[512] = {0};

// Device address of this device's copy of `scratch`, looked up once per
// device (with that device current).  The array's name in host code is only
// a stand-in; kernels get this address.
int scratch_address(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return SSG_ERR_HIP;
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SSG_OK;
}

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// bytes of scratch an nx x ny FP64 tendency needs (the layout below)
size_t scratch_need(int nx, int ny) {
    const size_t c = sizeof(cpx<double>);
    const size_t planes = (size_t)(nx / 2 + 1) * ny * c;
    return align_up(nx * c) + align_up(ny * c) + align_up((size_t)nx * ny * c) + 2 * align_up(planes);
}

// tw[m] = exp(-2 pi i m / n), the argument formed as on the host
__global__ void k_twiddles(cpx<double>* tw, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double a = -2.0 * M_PI * (double)m / (double)n;
    tw[m] = cpx<double>{cos(a), sin(a)};
}

int noalloc_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_noalloc_err, sizeof(g_noalloc_err), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace

const char* noalloc_last_error() { return g_noalloc_err; }

}  // namespace ssg_impl

extern "C" int ssg_tendency_noalloc(const ssg_domain_desc* desc, const double* b, const double* phi, double* out,
                                    int32_t ab2_reference_aliasing, void* stream) {
    using namespace ssg_impl;
    g_noalloc_err[0] = 0;
    if (!desc) return noalloc_err(SSG_ERR_ARG, "null domain descriptor");
    if (!b || !phi || !out) return noalloc_err(SSG_ERR_ARG, "null field pointer");
    if (!pow2_in_range(desc->nx) || !pow2_in_range(desc->ny))
        return noalloc_err(SSG_ERR_UNSUPPORTED, "nx, ny must be powers of two in [8, 4096] (got %d, %d)", desc->nx,
                           desc->ny);
    if (!(desc->lx > 0) || !(desc->ly > 0)) return noalloc_err(SSG_ERR_ARG, "domain lengths must be > 0");
    if (desc->precision != SSG_FP64)
        return noalloc_err(SSG_ERR_UNSUPPORTED, "the allocation-free tendency is FP64 only");
    if (desc->device < 0 || desc->device >= kMaxDevices)
        return noalloc_err(SSG_ERR_ARG, "device ordinal %d out of range", desc->device);
    const size_t need = scratch_need(desc->nx, desc->ny);
    if (need > kNoallocScratchBytes)
        return noalloc_err(SSG_ERR_UNSUPPORTED,
                           "a %dx%d grid needs %zu bytes of scratch; the allocation-free path holds %zu "
                           "(grids up to 1024x1024 points)",
                           desc->nx, desc->ny, need, kNoallocScratchBytes);
    if (hipSetDevice(desc->device) != hipSuccess) return noalloc_err(SSG_ERR_HIP, "hipSetDevice(%d) failed", desc->device);
    unsigned char* base = nullptr;
    if (scratch_address(desc->device, &base)) return noalloc_err(SSG_ERR_HIP, "hipGetSymbolAddress(scratch) failed");

    // a host-side context that only describes the grid and points at the
    // scratch regions; it owns no device memory
    ssg_ctx c;
    c.nx = desc->nx;
    c.ny = desc->ny;
    c.nz = 1;
    c.lx = desc->lx;
    c.ly = desc->ly;
    c.lz = 1.0;
    c.device = desc->device;
    c.stream = static_cast<hipStream_t>(stream);
    c.g.nx = c.nx;
    c.g.ny = c.ny;
    c.g.pitch = c.ny;
    c.g.K = c.nx / 3;
    c.g.Ky = 2 * (c.ny / 2 + 1) / 3;
    c.g.dkx = 2.0 * M_PI / c.lx;
    c.g.dky = 2.0 * M_PI / c.ly;
    c.ncols_full = c.nx / 2 + 1;
    c.ncols_tend = c.g.K + 1;
    const size_t cs = sizeof(cpx<double>);
    const size_t planes = (size_t)c.ncols_full * c.ny * cs;
    unsigned char* p = base;
    auto take = [&](size_t bytes) {
        unsigned char* r = p;
        p += align_up(bytes);
        return reinterpret_cast<cpx<double>*>(r);
    };
    PathBuffers<double>& D = c.d;
    D.twx = take(c.nx * cs);
    D.twy = take(c.ny * cs);
    D.w = take((size_t)c.nx * c.ny * cs);
    D.Fx = take(planes);
    D.Fy = take(planes);
    D.Tp = D.Fx;  // the column passes run in place on their first input

    auto run = [&]() -> int {
        hipLaunchKernelGGL(k_twiddles, dim3((c.nx + 255) / 256), dim3(256), 0, c.stream, D.twx, c.nx);
        hipLaunchKernelGGL(k_twiddles, dim3((c.ny + 255) / 256), dim3(256), 0, c.stream, D.twy, c.ny);
        HIP_OK(&c, hipGetLastError());
        const double scale = 1.0 / ((double)c.nx * c.ny);
        // u = -∂yφₛ, v = ∂xφₛ (src/setICs.jl:161-192)
        int rc = launch_rows_fwd<double, double, SRC_PAIR>(&c, phi, nullptr, D.Fx, nullptr, c.ncols_full);
        if (rc) return rc;
        if ((rc = launch_cols<double, OP_VEL>(&c, D.Fx, nullptr, D.Fx, D.Fy, c.ncols_full, scale))) return rc;
        RowsInvArgs<double> wa{};
        wa.P0 = D.Fx;
        wa.P1 = D.Fy;
        wa.ncols_in = c.ncols_full;
        wa.w_out = D.w;
        if ((rc = launch_rows_inv<double, MODE_PLANES, EPI_STORE_W>(&c, wa))) return rc;
        // -M (Dx(u b) + Dy(v b)) (src/nonlinear.jl:22-36, src/transforms.jl:430-465)
        if ((rc = launch_rows_fwd<double, double, SRC_FLUX>(&c, b, D.w, D.Fx, D.Fy, c.ncols_tend))) return rc;
        if ((rc = tendency_cols<double>(&c, ab2_reference_aliasing ? 1 : 0))) return rc;
        RowsInvArgs<double> ta{};
        ta.P0 = D.Tp;
        ta.ncols_in = c.ncols_tend;
        ta.out_r = out;
        return launch_rows_inv<double, MODE_PAIR, EPI_STORE>(&c, ta);
    };
    const int rc = run();
    if (rc) return noalloc_err(rc, "%s", c.err);
    return SSG_OK;
}
