/*
 * g(r) for callers that cannot allocate device memory: edmd_pcf_noalloc.
 *
 * The same pair histogram as edmd_pcf (k_pcf, calculate_pcf of pcf.c:16-75),
 * with positions and results in device memory the caller owns and the
 * integer pair counts in a fixed device array of this library, `scratch`,
 * instead of a buffer allocated per call.  Nothing here calls hipMalloc,
 * hipFree or an allocating routine; the caller passes inputs and outputs.
 *
 * `scratch` is a device symbol: every GPU the library runs on holds its own
 * copy at its own address for the life of the process.  Host code never
 * hands the symbol itself to a kernel (on the host it names a host-side
 * stand-in, not device memory); the address of the current device's copy is
 * looked up once per device with hipGetSymbolAddress and kept in `dev_ptr`.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several host threads, or a call on another stream while one is in
 * flight) race on it.  Each call clears the counts it uses before counting,
 * with a fill kernel.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/edmd.h"
#include "edmd_internal.h"

int edmd_internal_fail(int code, const char* msg);

/* Largest bin count of one call: 32768 bins of 8-byte counts, 256 KiB per
 * GPU (the bin limit of edmd_pcf). */
#define PCF_SCRATCH_BINS 32768
static_assert(PCF_SCRATCH_BINS == EDMD_PCF_NOALLOC_MAX_BINS, "header limit");

static __device__ __align__(256) unsigned long long scratch[PCF_SCRATCH_BINS];

#define PCF_MAX_DEVICES 64
static unsigned long long* dev_ptr[PCF_MAX_DEVICES];
static std::mutex dev_ptr_lock;

__global__ __launch_bounds__(256) void k_zero_counts(unsigned long long* __restrict__ cnt, int nb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) cnt[b] = 0ull;
}

/* The normalisation of calculate_pcf (pcf.c:16-75), in the order edmd_pcf's
 * host code evaluates it: r = (b + 1/2) dr, g = 2 count / (2 pi r dr *
 * (n / (Lx Ly)) * n). */
__global__ __launch_bounds__(256) void k_pcf_normalise(const unsigned long long* __restrict__ cnt,
                                                       int nb, int n, double Lx, double Ly,
                                                       double dr, double* __restrict__ r,
                                                       double* __restrict__ g) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double pi = 3.14159265358979323846;
    const double volume = Lx * Ly;
    const double density = n / volume;
    const double rb = (b + 0.5) * dr;
    const double shell = 2 * pi * rb * dr;
    const double norm = shell * density * n;
    const double c = 2.0 * (double)cnt[b];
    r[b] = rb;
    g[b] = norm > 0 ? c / norm : 0.0;
}

/* Address of the current device's copy of `scratch`, looked up on the first
 * call on that device. */
static int scratch_address(unsigned long long** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0)
        return edmd_internal_fail(EDMD_ERR_NODEVICE, "g(r): no current HIP device");
    if (dev >= PCF_MAX_DEVICES)
        return edmd_internal_fail(EDMD_ERR_ARG, "g(r): device index beyond 63");
    std::lock_guard<std::mutex> g(dev_ptr_lock);
    if (!dev_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return edmd_internal_fail(EDMD_ERR_HIP, "g(r): hipGetSymbolAddress of the scratch "
                                                    "counts failed");
        dev_ptr[dev] = (unsigned long long*)p;
    }
    *out = dev_ptr[dev];
    return EDMD_OK;
}

extern "C" int edmd_pcf_noalloc(int n, const double* x, const double* y, double Lx, double Ly,
                                double dr, double max_r, double* r, double* g, int* nbins) {
    if (n < 1 || !x || !y || !r || !g || !nbins || !(dr > 0) || !(max_r > 0) || !(Lx > 0) ||
        !(Ly > 0))
        return edmd_internal_fail(EDMD_ERR_ARG, "bad g(r) arguments");
    if (n > (1 << 22))
        return edmd_internal_fail(EDMD_ERR_ARG, "g(r): more than 2^22 positions");
    const double nbd = max_r / dr;
    if (!(nbd < (double)PCF_SCRATCH_BINS + 1))
        return edmd_internal_fail(EDMD_ERR_ARG, "g(r) without allocation takes at most 32768 bins");
    const int nb = (int)nbd;
    if (nb < 1) return edmd_internal_fail(EDMD_ERR_ARG, "g(r) needs at least one bin");
    if (*nbins < nb) {
        *nbins = nb;
        return edmd_internal_fail(EDMD_ERR_ARG, "g(r) output arrays are too short");
    }
    unsigned long long* cnt = nullptr;
    int rc = scratch_address(&cnt);
    if (rc) return rc;
    const hipStream_t s = nullptr;
    const int blocks = (nb + 255) / 256;
    hipLaunchKernelGGL(k_zero_counts, dim3(blocks), dim3(256), 0, s, cnt, nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = edmd_launch_pcf(x, y, nullptr, n, 0, 0, Lx, Ly, dr, nb, cnt, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_pcf_normalise, dim3(blocks), dim3(256), 0, s, cnt, nb, n, Lx, Ly, dr,
                           r, g);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return edmd_internal_fail(EDMD_ERR_HIP, hipGetErrorString(e));
    *nbins = nb;
    return EDMD_OK;
}
