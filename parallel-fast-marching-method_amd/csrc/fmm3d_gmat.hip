// fmm3d_gmat.hip -- the Frechet matrix G resident on the device and the
// sparse products subinv's subspace step takes with it (subinv.f90:
// 612-620 G^T d, 759-790 G a / Cd then G^T, 829-850 the same for the
// projected matrix).
//
// subinv reads frechet.out into CSR (rows = traveltimes, 1-based columns =
// vertices, subinv.f90:346-399), builds the transpose column by column with
// the rows of a column in ascending order (subinv.f90:457-475), and sums in
// single precision, entry by entry in that order.  Here the CSR is uploaded
// once; the transpose comes from a stable radix sort of the entries by
// column (for rows with ascending columns, the order subinv builds); the
// products run one thread per column (G^T x) or per row (G x / cd), each a
// sequential float32 sum in subinv's order without contraction, so they
// are subinv's numbers.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <string>
#include <vector>

#include "../../include/fmm3d.h"

struct fmm3d_gmat {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t nrows = 0, ncols = 0, nnz = 0;
    int64_t* rp = nullptr;      // CSR
    int32_t* col = nullptr;     // 0-based columns
    float* val = nullptr;
    int64_t* cp = nullptr;      // CSC: ncols+1 offsets
    int32_t* row = nullptr;     // rows ascending within a column
    float* tval = nullptr;
    float* x = nullptr;         // staging of one product's input and output
    float* y = nullptr;
    float* cd = nullptr;
    std::string err;
};

namespace {

thread_local std::string g_gmat_err;

int gfail(fmm3d_gmat* m, int code, const std::string& msg)
{
    if (m) m->err = msg;
    else g_gmat_err = msg;
    return code;
}

// the row of every CSR entry
__global__ void k_entry_rows(int64_t nrows, const int64_t* __restrict__ rp, int32_t* __restrict__ erow)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x)
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) erow[k] = (int32_t)r;
}

// column offsets of the column-sorted entries: cp[c] = first entry of column c
__global__ void k_col_offsets(int64_t nnz, int64_t ncols, const int32_t* __restrict__ scol, int64_t* __restrict__ cp)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k <= nnz; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c0 = k == 0 ? -1 : scol[k - 1];
        const int64_t c1 = k == nnz ? ncols : scol[k];
        for (int64_t c = c0 + 1; c <= c1; ++c) cp[c] = k;
    }
}

__global__ void k_gather_t(int64_t nnz, const int32_t* __restrict__ perm, const int32_t* __restrict__ erow,
                           const float* __restrict__ val, int32_t* __restrict__ row, float* __restrict__ tval)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = perm[k];
        row[k] = erow[e];
        tval[k] = val[e];
    }
}

// y(c) = sum over column c of G(r, c) x(r), rows ascending (subinv.f90:612-616)
__global__ void k_gtx(int64_t ncols, const int64_t* __restrict__ cp, const int32_t* __restrict__ row,
                      const float* __restrict__ tval, const float* __restrict__ x, float* __restrict__ y)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncols; c += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int64_t k = cp[c]; k < cp[c + 1]; ++k) s = s + tval[k] * x[row[k]];
        y[c] = s;
    }
}

// y(r) = (sum over row r of G(r, c) x(c)) / cd(r), in file order (subinv.f90:759-767)
__global__ void k_gx(int64_t nrows, const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                     const float* __restrict__ val, const float* __restrict__ x, const float* __restrict__ cd,
                     float* __restrict__ y)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) s = s + val[k] * x[col[k]];
        y[r] = cd ? s / cd[r] : s;
    }
}

int grid_of(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 65536); }

}  // namespace

extern "C" const char* fmm3d_gmat_last_error(const fmm3d_gmat* m) { return m ? m->err.c_str() : g_gmat_err.c_str(); }

extern "C" int fmm3d_gmat_create(int device, int64_t nrows, int64_t ncols, const int64_t* row_ptr, const int32_t* cols,
                                 const float* vals, fmm3d_gmat** out)
{
    if (!out || nrows < 0 || ncols <= 0 || !row_ptr) return gfail(nullptr, FMM3D_E_ARG, "fmm3d_gmat_create: bad arguments");
    *out = nullptr;
    const int64_t nnz = row_ptr[nrows];
    if (row_ptr[0] != 0 || nnz < 0 || nnz > INT32_MAX || ncols > INT32_MAX || nrows > INT32_MAX)
        return gfail(nullptr, FMM3D_E_ARG, "fmm3d_gmat_create: row_ptr must start at 0 and hold < 2^31 entries");
    if (nnz > 0 && (!cols || !vals)) return gfail(nullptr, FMM3D_E_ARG, "fmm3d_gmat_create: null entries");
    std::vector<int32_t> c0((size_t)nnz);
    for (int64_t r = 0; r < nrows; ++r) {
        if (row_ptr[r + 1] < row_ptr[r]) return gfail(nullptr, FMM3D_E_ARG, "fmm3d_gmat_create: row_ptr decreases");
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            if (cols[k] < 1 || cols[k] > ncols)
                return gfail(nullptr, FMM3D_E_ARG, "fmm3d_gmat_create: column " + std::to_string(cols[k]) + " out of 1.." +
                                                       std::to_string(ncols));
            if (k > row_ptr[r] && cols[k] <= cols[k - 1])
                return gfail(nullptr, FMM3D_E_ARG, "fmm3d_gmat_create: columns must ascend within a row");
            c0[k] = cols[k] - 1;
        }
    }
    if (hipSetDevice(device) != hipSuccess) return gfail(nullptr, FMM3D_E_DEVICE, "fmm3d_gmat_create: hipSetDevice failed");
    fmm3d_gmat* m = new fmm3d_gmat();
    m->device = device;
    m->nrows = nrows; m->ncols = ncols; m->nnz = nnz;
    int32_t *erow = nullptr, *iota = nullptr, *perm = nullptr, *scol = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    const size_t nz = (size_t)std::max<int64_t>(nnz, 1);
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&m->rp, (nrows + 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&m->col, nz * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&m->val, nz * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->cp, (ncols + 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&m->row, nz * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&m->tval, nz * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->x, std::max<int64_t>(nrows, ncols) * sizeof(float) + 256);
    if (e == hipSuccess) e = hipMalloc(&m->y, std::max<int64_t>(nrows, ncols) * sizeof(float) + 256);
    if (e == hipSuccess) e = hipMalloc(&m->cd, std::max<int64_t>(nrows, 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&erow, nz * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&iota, nz * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&perm, nz * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&scol, nz * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpyAsync(m->rp, row_ptr, (nrows + 1) * sizeof(int64_t), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(m->col, c0.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(m->val, vals, nnz * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess && nnz) {
        std::vector<int32_t> h((size_t)nnz);
        for (int64_t k = 0; k < nnz; ++k) h[k] = (int32_t)k;
        e = hipMemcpyAsync(iota, h.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);   // h is pageable and goes out of scope
    }
    if (e == hipSuccess && nnz) {
        k_entry_rows<<<grid_of(nrows), 256, 0, m->stream>>>(nrows, m->rp, erow);
        e = hipGetLastError();
    }
    // stable sort by column: within a column the entries keep their CSR (row) order
    int bits = 1;
    while (bits < 32 && (int64_t(1) << bits) < ncols) ++bits;
    if (e == hipSuccess && nnz)
        e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, m->col, scol, iota, perm, (size_t)nnz, 0, bits, m->stream);
    if (e == hipSuccess && nnz) e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1));
    if (e == hipSuccess && nnz)
        e = rocprim::radix_sort_pairs(tmp, tmp_bytes, m->col, scol, iota, perm, (size_t)nnz, 0, bits, m->stream);
    if (e == hipSuccess) {
        k_col_offsets<<<grid_of(nnz + 1), 256, 0, m->stream>>>(nnz, ncols, scol, m->cp);
        e = hipGetLastError();
    }
    if (e == hipSuccess && nnz) {
        k_gather_t<<<grid_of(nnz), 256, 0, m->stream>>>(nnz, perm, erow, m->val, m->row, m->tval);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(erow); (void)hipFree(iota); (void)hipFree(perm); (void)hipFree(scol); (void)hipFree(tmp);
    if (e != hipSuccess) {
        std::string msg = std::string("fmm3d_gmat_create: ") + hipGetErrorString(e);
        fmm3d_gmat_destroy(m);
        return gfail(nullptr, e == hipErrorOutOfMemory ? FMM3D_E_NOMEM : FMM3D_E_DEVICE, msg);
    }
    *out = m;
    return FMM3D_OK;
}

extern "C" int fmm3d_gmat_destroy(fmm3d_gmat* m)
{
    if (!m) return FMM3D_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->rp); (void)hipFree(m->col); (void)hipFree(m->val); (void)hipFree(m->cp); (void)hipFree(m->row);
    (void)hipFree(m->tval); (void)hipFree(m->x); (void)hipFree(m->y); (void)hipFree(m->cd);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return FMM3D_OK;
}

extern "C" int fmm3d_gmat_gtx(fmm3d_gmat* m, const float* x, float* y)
{
    if (!m || !x || !y) return gfail(m, FMM3D_E_ARG, "fmm3d_gmat_gtx: null argument");
    hipError_t e = hipSetDevice(m->device);
    if (e == hipSuccess && m->nrows) e = hipMemcpyAsync(m->x, x, m->nrows * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        k_gtx<<<grid_of(m->ncols), 256, 0, m->stream>>>(m->ncols, m->cp, m->row, m->tval, m->x, m->y);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, m->y, m->ncols * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return gfail(m, FMM3D_E_DEVICE, std::string("fmm3d_gmat_gtx: ") + hipGetErrorString(e));
    return FMM3D_OK;
}

extern "C" int fmm3d_gmat_gx(fmm3d_gmat* m, const float* x, const float* cd, float* y)
{
    if (!m || !x || !y) return gfail(m, FMM3D_E_ARG, "fmm3d_gmat_gx: null argument");
    hipError_t e = hipSetDevice(m->device);
    if (e == hipSuccess) e = hipMemcpyAsync(m->x, x, m->ncols * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess && cd && m->nrows) e = hipMemcpyAsync(m->cd, cd, m->nrows * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess && m->nrows) {
        k_gx<<<grid_of(m->nrows), 256, 0, m->stream>>>(m->nrows, m->rp, m->col, m->val, m->x, cd ? m->cd : nullptr, m->y);
        e = hipGetLastError();
    }
    if (e == hipSuccess && m->nrows) e = hipMemcpyAsync(y, m->y, m->nrows * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return gfail(m, FMM3D_E_DEVICE, std::string("fmm3d_gmat_gx: ") + hipGetErrorString(e));
    return FMM3D_OK;
}
