// Field probes on the device (include/saras_probes.h): probes::getData (lib/io/probes.cc:145-157)
// as one gather launch over (probe, field) pairs and one transfer of the values to the host.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/saras_mg.h"
#include "../../include/saras_probes.h"
#include "mg_state.h"

using namespace sr;

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            return set_err(SR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

namespace {

// the probed fields: device pointers and core sizes, passed by value to the gather kernel
struct ProbeFields {
    const double *f[SR_PROBE_MAX_FIELDS];
    int n[SR_PROBE_MAX_FIELDS][3];
};

// one thread per (probe, field): out[p * nf + f] = F_f(i, j, k) of the padded array
__global__ void __launch_bounds__(256) k_probe_gather(ProbeFields P, int nf, int np, const int *__restrict__ idx,
                                                      double *__restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np * nf) return;
    int p = t / nf, f = t - p * nf;
    int i = idx[3 * p], j = idx[3 * p + 1], k = idx[3 * p + 2];
    const int *n = P.n[f];
    long q = ((long)(i + 1) * (n[1] + 2) + (j + 1)) * (n[2] + 2) + (k + 1);
    out[t] = P.f[f][q];
}

int check_probe_args(int nfields, const double *const *fields, const int *dims, int nprobes, const int *idx,
                     const double *out, ProbeFields &P) {
    if (nfields < 1 || nfields > SR_PROBE_MAX_FIELDS) return set_err(SR_ERR_ARG, "nfields must be 1..10");
    if (nprobes < 0 || !fields || !dims || (nprobes > 0 && (!idx || !out))) return set_err(SR_ERR_ARG, "bad arguments");
    for (int f = 0; f < nfields; f++) {
        if (!fields[f]) return set_err(SR_ERR_ARG, "null field");
        for (int a = 0; a < 3; a++) {
            if (dims[3 * f + a] < 1) return set_err(SR_ERR_ARG, "bad field sizes");
            P.n[f][a] = dims[3 * f + a];
        }
        P.f[f] = fields[f];
    }
    for (int p = 0; p < nprobes; p++)
        for (int f = 0; f < nfields; f++)
            for (int a = 0; a < 3; a++)
                if (idx[3 * p + a] < -1 || idx[3 * p + a] > P.n[f][a])
                    return set_err(SR_ERR_ARG, "probe " + std::to_string(p) + " lies outside the padded range of field " +
                                                   std::to_string(f));
    return SR_OK;
}

}  // namespace

// Staging of sr_probe_gather_noalloc: the probe list (3 ints per probe) followed, at a 256-byte
// boundary, by the gathered values (SR_PROBE_MAX_FIELDS doubles per probe).  Every GPU holds its
// own copy for the life of the process; every call on a GPU shares it.
static constexpr size_t kScratchIdxBytes = ((size_t)SR_PROBE_NOALLOC_MAX * 3 * sizeof(int) + 255) / 256 * 256;
static constexpr size_t kScratchBytes = kScratchIdxBytes + (size_t)SR_PROBE_NOALLOC_MAX * SR_PROBE_MAX_FIELDS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// device address of `scratch` per device, taken once per device with that device current (the
// array's name is a host-side stand-in and is never used as an address on the host)
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

static int scratch_ptr(unsigned char **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_err(SR_ERR_STATE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err(SR_ERR_HIP, "no device address for the probe scratch");
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return SR_OK;
}

extern "C" int sr_probe_gather_noalloc(int nfields, const double *const *fields, const int *dims, int nprobes,
                                       const int *idx, double *out) {
    ProbeFields P{};
    int r = check_probe_args(nfields, fields, dims, nprobes, idx, out, P);
    if (r != SR_OK) return r;
    if (nprobes > SR_PROBE_NOALLOC_MAX)
        return set_err(SR_ERR_ARG, "more than SR_PROBE_NOALLOC_MAX probes for the allocation-free gather");
    if (nprobes == 0) return SR_OK;
    unsigned char *base = nullptr;
    r = scratch_ptr(&base);
    if (r != SR_OK) return r;
    int *didx = (int *)base;
    double *dout = (double *)(base + kScratchIdxBytes);
    const int total = nprobes * nfields;
    // every entry the kernel reads is written first (the probe list) and every value read back is
    // written by the kernel, so nothing has to start at zero
    HIPCHK(hipMemcpy(didx, idx, sizeof(int) * 3 * (size_t)nprobes, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_gather, dim3((total + 255) / 256), dim3(256), 0, 0, P, nfields, nprobes, didx, dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dout, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost));
    return SR_OK;
}

extern "C" int sr_probe_gather(int nfields, const double *const *fields, const int *dims, int nprobes, const int *idx,
                               double *out) {
    ProbeFields P{};
    int r = check_probe_args(nfields, fields, dims, nprobes, idx, out, P);
    if (r != SR_OK) return r;
    if (nprobes == 0) return SR_OK;
    int *didx = nullptr;
    double *dout = nullptr;
    const int total = nprobes * nfields;
    auto release = [&]() {
        if (didx) (void)hipFree(didx);
        if (dout) (void)hipFree(dout);
    };
    hipError_t e = hipMalloc((void **)&didx, sizeof(int) * 3 * (size_t)nprobes);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, sizeof(double) * (size_t)total);
    if (e == hipSuccess) e = hipMemcpy(didx, idx, sizeof(int) * 3 * (size_t)nprobes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_probe_gather, dim3((total + 255) / 256), dim3(256), 0, 0, P, nfields, nprobes, didx, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost);
    release();
    if (e != hipSuccess) return set_err(SR_ERR_HIP, std::string("probe gather: ") + hipGetErrorString(e));
    return SR_OK;
}
