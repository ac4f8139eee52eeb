// da_pairwise_distances_device_noalloc_s: the single-precision pairwise distances of
// device-resident X and Y into a device-resident D without any allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// The work is that of da_pairwise_distances_s (the same argument checks, metric switch,
// distance kernels and results); the device memory the host entry point allocates for
// row-major compact copies of the operands and for the row norms comes here from one
// file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES per GPU. Host code never
// passes the array's name to a kernel or a copy (on the host it is only a stand-in for the
// symbol): the array's device address is looked up once per device with
// hipGetSymbolAddress and that address is used instead. Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (from several host threads) race; calls
// from one thread are ordered on the null stream and each returns after its work is done.
#include "../metrics/pairwise_common.hpp"
#include "launch.hpp"
#include <hip/hip_runtime.h>
#include <limits>
#include <mutex>

namespace dagpu {

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex scratch_mu;
void *scratch_ptr[MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
hipError_t scratch_for_current_device(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        if (!p)
            return hipErrorInvalidSymbol;
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Whether p is device (or managed) memory of the current device's address space.
bool is_device_pointer(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

#define NA_HIP(call)                                                                         \
    do {                                                                                     \
        if ((call) != hipSuccess) {                                                          \
            (void)hipGetLastError();                                                         \
            return da_status_internal_error;                                                 \
        }                                                                                    \
    } while (0)

template <typename T>
da_status pairwise_noalloc(da_order order, da_int m, da_int n, da_int k, const T *X,
                           da_int ldx, const T *Y, da_int ldy, T *D, da_int ldd, T p,
                           da_metric metric) {
    da_status chk = pairwise_check<T>(order, m, n, k, X, ldx, Y, ldy, D, ldd, p, metric);
    if (chk != da_status_success)
        return chk;
    const int km = pairwise_kmetric<T>(metric, p);
    if (km < 0)
        return da_status_not_implemented;
    if (!is_device_pointer(X) || !is_device_pointer(D) || (Y && !is_device_pointer(Y)))
        return da_status_invalid_pointer;
    const bool x_is_y = (Y == nullptr);
    if (x_is_y) {
        n = m;
        ldy = ldx;
    }
    const bool colmaj = order == column_major;
    // Operands are read as compact row-major rows: a copy is needed for column-major input
    // and for row-major input with a leading dimension above k.
    const bool copy_x = colmaj || ldx != k;
    const bool copy_y = !x_is_y && (colmaj || ldy != k);
    const bool norms = pairwise_needs_norms(km);
    const size_t bx = copy_x ? align256((size_t)m * k * sizeof(T)) : 0;
    const size_t by = copy_y ? align256((size_t)n * k * sizeof(T)) : 0;
    const size_t bnx = norms ? align256((size_t)m * sizeof(double)) : 0;
    const size_t bny = (norms && !x_is_y) ? align256((size_t)n * sizeof(double)) : 0;
    if (bx + by + bnx + bny > NOALLOC_SCRATCH_BYTES)
        return da_status_invalid_input; // more than the fixed scratch of 32 MiB per GPU

    void *base = nullptr;
    NA_HIP(scratch_for_current_device(&base));
    unsigned char *sp = static_cast<unsigned char *>(base);
    hipStream_t s = nullptr;

    auto compact = [&](const T *A, da_int rows, da_int ld, bool copy, size_t off,
                       const T **out) -> hipError_t {
        if (!copy) {
            *out = A;
            return hipSuccess;
        }
        T *dst = reinterpret_cast<T *>(sp + off);
        *out = dst;
        return colmaj ? launch_colmajor_to_rowmajor<T>(A, ld, rows, k, dst, s)
                      : launch_pack_rows<T, T>(A, ld, rows, k, dst, k, rows, s);
    };
    const T *dX = nullptr, *dY = nullptr;
    NA_HIP(compact(X, m, ldx, copy_x, 0, &dX));
    if (x_is_y)
        dY = dX;
    else
        NA_HIP(compact(Y, n, ldy, copy_y, bx, &dY));
    double *dnx = nullptr, *dny = nullptr;
    if (norms) {
        dnx = reinterpret_cast<double *>(sp + bx + by);
        NA_HIP(launch_row_sqnorms<T>(dX, k, m, k, dnx, s));
        if (x_is_y) {
            dny = dnx;
        } else {
            dny = reinterpret_cast<double *>(sp + bx + by + bnx);
            NA_HIP(launch_row_sqnorms<T>(dY, k, n, k, dny, s));
        }
        if (km == KM_COSINE) {
            const double eps = std::numeric_limits<T>::epsilon();
            NA_HIP(launch_cosine_norms(dnx, m, eps, s));
            if (!x_is_y)
                NA_HIP(launch_cosine_norms(dny, n, eps, s));
        }
    }
    // Row-major output: rows of X against Y. Column-major output: the transposed problem,
    // rows of Y against X written row by row, which is D in column-major order with ldd.
    const T *A = colmaj ? dY : dX;
    const T *B = colmaj ? dX : dY;
    const double *na = colmaj ? dny : dnx;
    const double *nb = colmaj ? dnx : dny;
    const da_int ma = colmaj ? n : m;
    const da_int ncols = colmaj ? m : n;
    const int zero_diag = pairwise_zero_diag(km, x_is_y);
    // At most 65,535 row tiles of 64 per launch (the grid's y extent).
    const da_int chunk = 65535 * 64;
    for (da_int r0 = 0; r0 < ma; r0 += chunk) {
        const da_int rows = ma - r0 < chunk ? ma - r0 : chunk;
        NA_HIP(pairwise_tile<T>(km, A + (size_t)r0 * k, B, rows, ncols, k, D + (size_t)r0 * ldd,
                                ldd, na ? na + r0 : nullptr, nb, (double)p, zero_diag, (int)r0,
                                s));
    }
    NA_HIP(hipStreamSynchronize(s));
    return da_status_success;
}

} // namespace
} // namespace dagpu

extern "C" {

da_status da_pairwise_distances_device_noalloc_s(da_order order, da_int m, da_int n, da_int k,
                                                 const float *d_X, da_int ldx,
                                                 const float *d_Y, da_int ldy, float *d_D,
                                                 da_int ldd, float p, da_metric metric) {
    return dagpu::pairwise_noalloc<float>(order, m, n, k, d_X, ldx, d_Y, ldy, d_D, ldd, p,
                                          metric);
}

} // extern "C"
