// Side-by-side timing of rocFFT (through hipFFT, used as a black box) for
// the spectral round trip the library performs per filter application: a 2D
// real-to-complex transform, a diagonal multiply, and the complex-to-real
// transform back.  Built as its own library (libssg_rocfft_cmp.so) so the
// product library does not link hipFFT.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdint>

namespace {

template <typename C>
__global__ void k_scale_spec(C* s, long n, double f) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        s[q].x *= f;
        s[q].y *= f;
    }
}

}  // namespace

extern "C" {

// Times `reps` back-to-back round trips R2C -> scale -> C2R of an (ny, nx)
// x-fastest real field with HIP events; us_out[0] = microseconds per round
// trip, us_out[1] = per forward transform, us_out[2] = per inverse.
// precision: 0 = FP64 (D2Z/Z2D), 1 = FP32 (R2C/C2R).  Returns 0 on success.
int ssg_cmp_rocfft(int nx, int ny, int precision, int reps, double* us_out) {
    if (nx < 2 || ny < 2 || reps < 1 || !us_out) return -1;
    const size_t nreal = (size_t)nx * ny, ncplx = (size_t)ny * (nx / 2 + 1);
    const size_t rs = precision ? sizeof(float) : sizeof(double);
    void *r = nullptr, *c = nullptr;
    if (hipMalloc(&r, nreal * rs) != hipSuccess) return -2;
    if (hipMalloc(&c, ncplx * 2 * rs) != hipSuccess) {
        (void)hipFree(r);
        return -2;
    }
    (void)hipMemset(r, 0, nreal * rs);
    hipfftHandle pf, pi;
    const hipfftType tf = precision ? HIPFFT_R2C : HIPFFT_D2Z, ti = precision ? HIPFFT_C2R : HIPFFT_Z2D;
    int rc = 0;
    if (hipfftPlan2d(&pf, ny, nx, tf) != HIPFFT_SUCCESS || hipfftPlan2d(&pi, ny, nx, ti) != HIPFFT_SUCCESS) rc = -3;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto fwd = [&]() {
        return precision ? hipfftExecR2C(pf, (hipfftReal*)r, (hipfftComplex*)c)
                         : hipfftExecD2Z(pf, (hipfftDoubleReal*)r, (hipfftDoubleComplex*)c);
    };
    auto inv = [&]() {
        return precision ? hipfftExecC2R(pi, (hipfftComplex*)c, (hipfftReal*)r)
                         : hipfftExecZ2D(pi, (hipfftDoubleComplex*)c, (hipfftDoubleReal*)r);
    };
    auto scale = [&]() {
        const long n = (long)ncplx;
        const int blocks = (int)((n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256);
        if (precision)
            hipLaunchKernelGGL((k_scale_spec<float2>), dim3(blocks), dim3(256), 0, 0, (float2*)c, n, 1.0 / nreal);
        else
            hipLaunchKernelGGL((k_scale_spec<double2>), dim3(blocks), dim3(256), 0, 0, (double2*)c, n, 1.0 / nreal);
    };
    float ms = 0;
    if (rc == 0) {
        for (int i = 0; i < 3; ++i) {  // warm-up (plans, caches)
            fwd();
            scale();
            inv();
        }
        (void)hipEventRecord(e0, 0);
        for (int i = 0; i < reps; ++i) {
            fwd();
            scale();
            inv();
        }
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        us_out[0] = 1e3 * ms / reps;
        (void)hipEventRecord(e0, 0);
        for (int i = 0; i < reps; ++i) fwd();
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        us_out[1] = 1e3 * ms / reps;
        (void)hipEventRecord(e0, 0);
        for (int i = 0; i < reps; ++i) inv();
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        us_out[2] = 1e3 * ms / reps;
        if (hipDeviceSynchronize() != hipSuccess) rc = -4;
        hipfftDestroy(pf);
        hipfftDestroy(pi);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(r);
    (void)hipFree(c);
    return rc;
}

}  // extern "C"
