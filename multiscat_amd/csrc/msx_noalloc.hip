// Allocation-free entry point: the symmetric eigensystem of msx_symmetric_eigensystem on device
// memory the caller owns, with every intermediate in one file-scope device array.
//
// The array (`scratch`, about 10 MiB) exists once per GPU the library runs on, for the life of
// the process. Host code never passes the array's name to a kernel or a copy: on the host the
// name is only a stand-in for the symbol. The device address is looked up once per device with
// hipGetSymbolAddress (that device current) and kept; every use goes through it.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's contents: calls that
// overlap on one GPU (two streams, two threads) race and must be serialised by the caller.
// Nothing in the array needs clearing: the Householder stage, the divide-and-conquer tree and
// the back-transformation write every value they read.
#include <mutex>

#include "../../include/multiscat_amd.h"
#include "msx_internal.h"

using namespace msx;

namespace {

constexpr int kNoallocMaxN = MSX_NOALLOC_EIG_MAX_N;
// A's working copy (n^2), the divide-and-conquer workspace (3 n^2 + 4 n) and Q^T (n^2)
constexpr size_t kScratchDoubles =
    5ull * kNoallocMaxN * kNoallocMaxN + 4ull * kNoallocMaxN;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up on the first call on the device).
double *scratch_address() {
    int dev = 0;
    MSX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(MSX_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        MSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(MSX_ERR_DEVICE, "scratch array has no device address");
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    return g_scratch_dev[dev];
}

int fail_noalloc(const Error &e) {
    set_last_error(e.what());
    return e.code;
}

}  // namespace

extern "C" int msx_symmetric_eigensystem_noalloc(int n, const double *A, double *lam, double *Q,
                                                 int *info, void *stream) {
    try {
        if (!A || !lam || !Q || !info) throw Error(MSX_ERR_INVALID, "null argument");
        if (n < 2 || n > kNoallocMaxN)
            throw Error(MSX_ERR_DIMS, "msx_symmetric_eigensystem_noalloc: 2 <= n <= " +
                                          std::to_string(kNoallocMaxN) + " required, got n = " +
                                          std::to_string(n));
        hipStream_t s = static_cast<hipStream_t>(stream);
        double *base = scratch_address();
        const size_t n2 = (size_t)n * n;
        double *A_work = base;
        double *work = A_work + n2;
        double *QT = work + eig_dc_workspace(n);
        // the solver destroys its matrix: work on a copy, A stays the caller's
        MSX_HIP(hipMemcpyAsync(A_work, A, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
        launch_eigensystem_dc(n, A_work, work, lam, Q, QT, info, nullptr, s);
        return MSX_OK;
    } catch (const Error &e) {
        return fail_noalloc(e);
    }
}
