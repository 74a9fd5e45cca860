// api_noalloc.cpp -- ipg_landau_noalloc: the Landau frame and the spatial
// eccentricities of a T^{mu nu} that is already in device memory, with no
// device allocation.
//
// ipg_landau (the context entry point) runs k_landau ... k_ecc_finish on the
// event's own device block.  The kernels need, beyond their inputs (T, the
// colour-charge densities) and outputs (the Landau frame, the observables),
// a descriptor, per-block partial sums, a per-site flag and running-eps
// plane and the observable slots.  Here those live in one file-scope device
// array, `scratch`, whose address is looked up once per device with
// hipGetSymbolAddress; every call on a device reuses that device's copy.
// Calls that overlap on one device race on it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "host.hpp"
#include "ipg_internal.hpp"

namespace {

// largest lattice the library supports (fft_lds.hpp); the scratch layout is
// sized for it
constexpr int kMaxN = 1024;
constexpr long long kMaxS = (long long)kMaxN * kMaxN;
constexpr int kMaxBlocks = (int)((((kMaxS + IPG_BLOCK - 1) / IPG_BLOCK) + 7) / 8 * 8);
constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }
// byte offsets: descriptor | observables | partials (32 slots per block) |
// running eps per site | flags per site
constexpr size_t kOffDesc = 0;
constexpr size_t kOffObs = align256(sizeof(EvDesc));
constexpr size_t kOffPartial = kOffObs + align256(sizeof(double) * kObsDoubles);
constexpr size_t kOffEpsv = kOffPartial + align256(sizeof(double) * 32 * (size_t)kMaxBlocks);
constexpr size_t kOffFlag = kOffEpsv + align256(sizeof(double) * (size_t)kMaxS);
constexpr size_t kScratchBytes = kOffFlag + align256(sizeof(int) * (size_t)kMaxS);  // 13.6 MB

__device__ __align__(256) unsigned char scratch[kScratchBytes];

int fail_na(int code, const std::string& msg) {
    ipg_record_error(msg);
    return code;
}

// scratch's device address, per device (each device holds its own copy)
constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {nullptr};

int scratch_ptr(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_na(IPG_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail_na(IPG_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail_na(IPG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = (unsigned char*)p;
    }
    *out = g_ptr[dev];
    return IPG_OK;
}

}  // namespace

extern "C" {

int ipg_landau_noalloc(const ipg_params* p, const double* T, const double* g2mu2A, const double* g2mu2B, int it,
                       double* land, ipg_observables* o) {
    if (!p || !T || !land || !o) return fail_na(IPG_ERR_ARG, "null argument");
    if (it < 1) return fail_na(IPG_ERR_ARG, "the Landau frame needs tau > 0 (it >= 1)");
    const int N = p->size;
    if (N < 5 || N > kMaxN)
        return fail_na(IPG_ERR_ARG, "lattice size " + std::to_string(N) + " outside [5, " + std::to_string(kMaxN) +
                                        "], the range the scratch array is sized for");
    unsigned char* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    ipg_params q = *p;
    ipg::derive(&q);
    EvDesc d;
    std::memset(&d, 0, sizeof(d));
    d.N = N;
    d.Ny = q.Ny;
    d.S = (long long)N * N;
    d.itmax = it;
    d.block0 = 0;
    d.nblocks = (int)(((d.S + IPG_BLOCK - 1) / IPG_BLOCK + 7) / 8 * 8);
    d.active = 1;
    d.g = q.g;
    d.L = q.L;
    d.a = q.L / N;
    d.dtau = q.dtau;
    d.QsmuRatio = q.QsmuRatio;
    d.QsmuRatioB = q.QsmuRatio;
    d.T = const_cast<double*>(T);
    d.g2A = const_cast<double*>(g2mu2A);
    d.g2B = const_cast<double*>(g2mu2B);
    d.land = land;
    d.obs = (double*)(base + kOffObs);
    d.partial = (double*)(base + kOffPartial);
    d.land_epsv = (double*)(base + kOffEpsv);
    d.land_flag = (int*)(base + kOffFlag);
    EvDesc* dd = (EvDesc*)(base + kOffDesc);
    hipStream_t st = nullptr;  // the default stream
    hipError_t e = hipMemcpy(dd, &d, sizeof(d), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(d.obs, 0, sizeof(double) * kObsDoubles, st);
    if (e != hipSuccess) return fail_na(IPG_ERR_DEVICE, std::string("scratch set-up: ") + hipGetErrorString(e));
    ipg_launch_landau(dd, 1, d.nblocks, it, 1, st);
    e = hipGetLastError();
    if (e != hipSuccess) return fail_na(IPG_ERR_DEVICE, std::string("Landau-frame launch: ") + hipGetErrorString(e));
    double h[kObsDoubles];
    e = hipMemcpyAsync(h, d.obs, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail_na(IPG_ERR_DEVICE, std::string("observables: ") + hipGetErrorString(e));
    ipg_obs_from_device(h, o);
    o->tau_fm = it * q.dtau * d.a;
    return IPG_OK;
}

}  // extern "C"
