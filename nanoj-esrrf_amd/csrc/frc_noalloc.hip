// frc_noalloc.hip -- the FRC of two device images without any allocation: the same kernels
// and sums as esrrf_frc_run_device, with every buffer carved out of one file-scope device
// array (`scratch`) instead of memory the handle allocates.  For callers that may not
// allocate device memory (serving loops, harnesses that forbid hipMalloc).
//
// The array lives once per GPU for the life of the process.  Its device address is taken
// with hipGetSymbolAddress the first time the entry point runs on a device, with that device
// current, and kept per device; host code never passes the symbol itself.  Calls that
// overlap on one GPU share the array and race: callers serialise them.  Every byte a call
// reads is written by the same call first (window and twiddles by k_frc_tables, the padded
// spectra by k_frc_prep, ring partials and sums by k_frc_rings / k_frc_reduce), so nothing
// needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/esrrf.h"
#include "frc_internal.h"

namespace esrrf {

namespace {

constexpr int kMaxN = ESRRF_FRC_NOALLOC_MAX_N;
constexpr int kParts = 512;   // ring partial sums per launch, as esrrf_frc_create uses

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kBufBytes = align256(sizeof(float2) * 2 * (size_t)kMaxN * kMaxN);
constexpr size_t kPartBytes = align256(sizeof(double) * 3 * (size_t)(kMaxN / 2) * kParts);
constexpr size_t kRingBytes = align256(sizeof(double) * 3 * (size_t)(kMaxN / 2));
constexpr size_t kTwBytes = align256(sizeof(float2) * (size_t)kMaxN);
constexpr size_t kWinBytes = align256(sizeof(float) * (size_t)kMaxN);
constexpr size_t kScratchBytes = kBufBytes + kPartBytes + kRingBytes + kTwBytes + kWinBytes;
static_assert(kScratchBytes <= 80u * 1024 * 1024, "scratch stays below 80 MiB");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// the Tukey(0.25) window of S points and the N twiddles e^{-2 pi i k / N}, in double and
// with the operations of esrrf_frc_create's host tables
__global__ void k_frc_tables(float *__restrict__ win, int S, float2 *__restrict__ tw, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) {
        const double alpha = 0.25;
        double v = 1.0;
        if (S > 1) {
            const double x = (double)i / (S - 1);
            if (x < alpha / 2) v = 0.5 * (1.0 + cos(2.0 * M_PI / alpha * (x - alpha / 2)));
            else if (x > 1.0 - alpha / 2)
                v = 0.5 * (1.0 + cos(2.0 * M_PI / alpha * (x - 1.0 + alpha / 2)));
        }
        win[i] = (float)v;
    }
    if (i < N) {
        const double ang = -2.0 * M_PI * i / N;
        tw[i] = make_float2((float)cos(ang), (float)sin(ang));
    }
}

// device address of `scratch` on the current device (taken once per device)
hipError_t scratch_of(int dev, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t st = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (st != hipSuccess) return st;
        g_scratch[dev] = (unsigned char *)ptr;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

}  // namespace esrrf

using namespace esrrf;

int esrrf_frc_run_device_noalloc(int device, void *hip_stream, const float *d_a,
                                 const float *d_b, int width, int height, double *resolution,
                                 float *curve, int *n_rings)
{
    if (!d_a || !d_b || width < 2 || height < 2) return ESRRF_ERR_INVALID;
    const int S = width < height ? width : height;
    int N = 2, lg = 1;
    while (N < S) {
        N <<= 1;
        lg++;
    }
    if (N > kMaxN) return ESRRF_ERR_UNSUPPORTED;   // the scratch holds transforms up to kMaxN
    int dev = device;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return ESRRF_ERR_HIP;
    if (dev >= kMaxDevices) return ESRRF_ERR_UNSUPPORTED;
    if (hipSetDevice(dev) != hipSuccess) return ESRRF_ERR_HIP;
    unsigned char *base = nullptr;
    if (scratch_of(dev, &base) != hipSuccess) {
        (void)hipGetLastError();
        return ESRRF_ERR_HIP;
    }
    hipStream_t s = (hipStream_t)hip_stream;
    FrcPlan p{};
    p.W = width;
    p.H = height;
    p.S = S;
    p.x0 = (width - S) / 2;
    p.y0 = (height - S) / 2;
    p.N = N;
    p.log2N = lg;
    p.nparts = kParts;
    p.buf = (float2 *)base;
    p.part = (double *)(base + kBufBytes);
    p.rings = (double *)(base + kBufBytes + kPartBytes);
    float2 *tw = (float2 *)(base + kBufBytes + kPartBytes + kRingBytes);
    float *win = (float *)(base + kBufBytes + kPartBytes + kRingBytes + kTwBytes);
    p.tw = tw;
    p.win = win;
    const int R = N / 2;
    double h_rings[3 * (kMaxN / 2)];
    hipLaunchKernelGGL(k_frc_tables, dim3((N + 255) / 256), dim3(256), 0, s, win, S, tw, N);
    hipError_t st = hipGetLastError();
    if (st == hipSuccess) st = launch_frc(p, d_a, d_b, s);
    if (st == hipSuccess)
        st = hipMemcpyAsync(h_rings, p.rings, sizeof(double) * 3 * R, hipMemcpyDeviceToHost, s);
    if (st == hipSuccess) st = hipStreamSynchronize(s);
    if (st != hipSuccess) {
        (void)hipGetLastError();
        return ESRRF_ERR_HIP;
    }
    const double res = frc_resolution(h_rings, N, curve);
    if (resolution) *resolution = res;
    if (n_rings) *n_rings = R;
    return ESRRF_OK;
}
