// Allocation-free conj_grad: the computation of npbcg_conj_grad_device
// (conj_grad_gpu, cg.cu:803-846) on caller-owned device arrays, with every
// buffer the kernels need beyond the caller's inputs and outputs (per-workgroup
// partial sums and the CG scalars) in a fixed-size device array of this
// library, so the call makes no device allocation.  For callers that cannot
// allocate: code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc.
//
// Kernels: the 64-lane CSR SpMV (spmv_variant 1) and the reference's step
// order (q = A p; alpha; z, r; rho; beta; p), the same launches, in the same
// order and with the same partial counts as a context created with
// spmv_variant = 1 and cg_form = 1, so the results are bit-identical to it.
//
// The array is shared by every call on one GPU and keeps the last call's
// contents; the scalars are cleared by a kernel at the start of each call
// (no hipMemsetAsync, so the sequence can be captured into a graph).  Calls
// that overlap on one GPU race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "cg_device.h"
#include "cg_kernels.h"
#include "npbcg.h"
#include "npbcg_internal.h"

using npbcg_detail::clear_error;
using npbcg_detail::set_error;

namespace {

// Partial sums: 7 vec_blocks + 4 spmv_blocks doubles (vec_blocks <= 8 CUs,
// spmv_blocks <= 4 CUs: 18,432 on 256 CUs), then the Scalars block.
constexpr size_t kScratchBytes = NPBCG_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU (each holds its own copy), taken
// once per device with that device current.
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_on(int dev, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(NPBCG_ERR_DEVICE,
                             std::string("npbcg_conj_grad_device_noalloc: hipGetSymbolAddress: ") +
                                 hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return NPBCG_OK;
}

__global__ void store_rnorm_kernel(const npbcg_dev::Scalars *sc, double *out) {
    if (threadIdx.x == 0) *out = sc->rnorm;
}

}  // namespace

extern "C" int npbcg_conj_grad_device_noalloc(int n, const int *rowstr, const int *colidx,
                                              const double *a, double *x, double *z, double *p,
                                              double *q, double *r, int cgitmax,
                                              double *rnorm, void *stream) {
    if (n < 1 || n > INT32_MAX - 64 || !rowstr || !colidx || !a || !x || !z || !p || !q || !r ||
        !rnorm || cgitmax < 1)
        return set_error(NPBCG_ERR_INVALID, "npbcg_conj_grad_device_noalloc: invalid argument");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return set_error(NPBCG_ERR_DEVICE, std::string("npbcg_conj_grad_device_noalloc: ") +
                                               hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(NPBCG_ERR_INVALID, "npbcg_conj_grad_device_noalloc: device ordinal " +
                                                std::to_string(dev) + " out of range");
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || cus <= 0) cus = 256;

    // launch shape of a context with spmv_variant 1 (npbcg_create)
    const int rpb = npbcg_dev::spmv_rows_per_block(npbcg_dev::kSpmvSeg64);
    const int vec_blocks = npbcg_dev::vector_blocks(n, cus);
    const int spmv_blocks = std::max(1, std::min((n + rpb - 1) / rpb, 4 * cus));
    const size_t part_doubles = 7 * (size_t)vec_blocks + 4 * (size_t)spmv_blocks;
    const size_t sc_off = (part_doubles * sizeof(double) + 255) & ~(size_t)255;
    const size_t need = sc_off + npbcg_dev::scalars_bytes();
    if (need > kScratchBytes)
        return set_error(NPBCG_ERR_CAPACITY,
                         "npbcg_conj_grad_device_noalloc: this launch shape needs " +
                             std::to_string(need) + " bytes of scratch, more than the " +
                             std::to_string(kScratchBytes) + " the library holds per GPU");
    void *base = nullptr;
    int rc = scratch_on(dev, &base);
    if (rc != NPBCG_OK) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);

    npbcg_dev::Fp64Problem pb{};
    pb.n = n;
    pb.rowstr = rowstr;
    pb.colidx = colidx;
    pb.a = a;
    pb.x = x;
    pb.z = z;
    pb.p = p;
    pb.q = q;
    pb.r = r;
    pb.p_alt = p;
    double *pp = static_cast<double *>(base);
    const size_t nv = vec_blocks, ns = spmv_blocks;
    pb.part.xx = pp;
    pb.part.rr[0] = pp + nv;
    pb.part.rr[1] = pp + 2 * nv;
    pb.part.nrm = pp + 3 * nv;
    pb.part.pq = pp + 5 * nv;
    pb.part.res = pp + 5 * nv + ns;
    pb.part.rz[0] = pp + 5 * nv + 4 * ns;
    pb.part.rz[1] = pp + 6 * nv + 4 * ns;
    pb.sc = reinterpret_cast<npbcg_dev::Scalars *>(static_cast<unsigned char *>(base) + sc_off);
    pb.vec_blocks = vec_blocks;
    pb.spmv_blocks = spmv_blocks;
    pb.spmv_variant = npbcg_dev::kSpmvSeg64;

    // scalars from zero, then the conj_grad bookkeeping (no stop test)
    const int sc_doubles = (int)(npbcg_dev::scalars_bytes() / sizeof(double));
    e = npbcg_dev::launch_fill(reinterpret_cast<double *>(pb.sc), sc_doubles, 0.0, 1, st);
    if (e == hipSuccess) e = npbcg_dev::launch_reset_scalars(pb, 0.0, st);
    if (e == hipSuccess) e = npbcg_dev::enqueue_conj_grad(pb, cgitmax, 0, 0, st, nullptr);
    if (e == hipSuccess) {
        store_rnorm_kernel<<<1, 64, 0, st>>>(pb.sc, rnorm);
        e = hipGetLastError();
    }
    if (e != hipSuccess)
        return set_error(NPBCG_ERR_DEVICE, std::string("npbcg_conj_grad_device_noalloc: ") +
                                               hipGetErrorString(e));
    clear_error();
    return NPBCG_OK;
}
