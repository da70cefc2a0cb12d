// d4_gpulu.hip -- the host simulation's linear solves as a dense LU on the GPU
// (rocSOLVER dgetrf / dgetrs, partial pivoting), in place of UMFPACK's
// umfpack_di_symbolic / _numeric / _solve (simulation.cpp:635-659).
//
// At the sizes of a realization (N <= a few thousand) one dense LU on the GPU
// takes milliseconds where a sparse LU of the cutoff pattern on a host core
// takes up to seconds (the 2D cutoff graph fills in heavily), so this is the
// solver the batched ensemble and large single runs use.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <map>
#include <mutex>
#include <thread>

#include "d4_solver.h"

namespace d4host {
namespace {

// one stream + rocBLAS handle per (thread, device); kept until the process ends
struct ThreadRes {
    hipStream_t stream = nullptr;
    rocblas_handle handle = nullptr;
};

ThreadRes* thread_res(int device)
{
    static std::mutex m;
    static std::map<std::pair<std::thread::id, int>, ThreadRes*> reg;
    std::lock_guard<std::mutex> l(m);
    auto key = std::make_pair(std::this_thread::get_id(), device);
    auto it = reg.find(key);
    if (it != reg.end()) return it->second;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    ThreadRes* r = new ThreadRes();
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess ||
        rocblas_create_handle(&r->handle) != rocblas_status_success ||
        rocblas_set_stream(r->handle, r->stream) != rocblas_status_success) {
        delete r;
        return nullptr;
    }
    reg[key] = r;
    return r;
}

__global__ void zero_kernel(double* A, size_t count)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) A[k] = 0.0;
}

// column c of the CSC matrix into column c of the dense column-major A
__global__ void scatter_kernel(const int* Ap, const int* Ai, const double* Ax, double* A, int n)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double* col = A + (size_t)c * n;
    for (int k = Ap[c]; k < Ap[c + 1]; ++k) col[Ai[k]] = Ax[k];
}

class GpuDenseLU : public LinearSolver {
public:
    explicit GpuDenseLU(int device) : device_(device) {}
    ~GpuDenseLU() override
    {
        for (void* p : {(void*)A_, (void*)ipiv_, (void*)info_, (void*)b_, (void*)Ap_, (void*)Ai_, (void*)Ax_})
            if (p) (void)hipFree(p);
    }

    int factor(int n, const int* Ap, const int* Ai, const double* Ax) override
    {
        singular_ = false;
        if (n <= 0 || !Ap) return n == 0 ? 0 : -1;
        ThreadRes* r = thread_res(device_);
        if (!r) return -2;
        const int nnz = Ap[n];
        if (grow(n, nnz)) return -2;
        n_ = n;
        const size_t nn = (size_t)n * n;
        hipError_t e = hipMemcpyAsync(Ap_, Ap, sizeof(int) * (n + 1), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess && nnz > 0)
            e = hipMemcpyAsync(Ai_, Ai, sizeof(int) * nnz, hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess && nnz > 0)
            e = hipMemcpyAsync(Ax_, Ax, sizeof(double) * nnz, hipMemcpyHostToDevice, r->stream);
        if (e != hipSuccess) return -2;
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, r->stream, A_, nn);
        hipLaunchKernelGGL(scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, r->stream, Ap_, Ai_, Ax_, A_, n);
        if (hipGetLastError() != hipSuccess) return -2;
        if (rocsolver_dgetrf(r->handle, n, n, A_, n, ipiv_, info_) != rocblas_status_success) return -2;
        int info = 0;
        if (hipMemcpyAsync(&info, info_, sizeof(int), hipMemcpyDeviceToHost, r->stream) != hipSuccess ||
            hipStreamSynchronize(r->stream) != hipSuccess)
            return -2;
        singular_ = info > 0;  // U(info, info) == 0: UMFPACK_WARNING_singular_matrix's case
        return singular_ ? 1 : 0;
    }

    int solve(const double* b, double* x) override
    {
        ThreadRes* r = thread_res(device_);
        if (!r || !A_) return -2;
        const int n = n_;
        if (hipMemcpyAsync(b_, b, sizeof(double) * n, hipMemcpyHostToDevice, r->stream) != hipSuccess) return -2;
        if (rocsolver_dgetrs(r->handle, rocblas_operation_none, n, 1, A_, n, ipiv_, b_, n) != rocblas_status_success)
            return -2;
        if (hipMemcpyAsync(x, b_, sizeof(double) * n, hipMemcpyDeviceToHost, r->stream) != hipSuccess ||
            hipStreamSynchronize(r->stream) != hipSuccess)
            return -2;
        return 0;
    }

    bool singular() const override { return singular_; }

private:
    int device_;
    int n_ = 0, cap_n_ = 0, cap_nnz_ = 0;
    bool singular_ = false;
    double* A_ = nullptr;
    int* ipiv_ = nullptr;
    int* info_ = nullptr;
    double* b_ = nullptr;
    int* Ap_ = nullptr;
    int* Ai_ = nullptr;
    double* Ax_ = nullptr;

    int grow(int n, int nnz)
    {
        if (hipSetDevice(device_) != hipSuccess) return -1;
        if (n > cap_n_) {
            for (void* p : {(void*)A_, (void*)ipiv_, (void*)info_, (void*)b_, (void*)Ap_})
                if (p) (void)hipFree(p);
            A_ = nullptr;
            ipiv_ = info_ = Ap_ = nullptr;
            b_ = nullptr;
            if (hipMalloc(&A_, sizeof(double) * (size_t)n * n) != hipSuccess ||
                hipMalloc(&ipiv_, sizeof(int) * n) != hipSuccess || hipMalloc(&info_, sizeof(int)) != hipSuccess ||
                hipMalloc(&b_, sizeof(double) * n) != hipSuccess ||
                hipMalloc(&Ap_, sizeof(int) * (n + 1)) != hipSuccess) {
                cap_n_ = 0;
                return -1;
            }
            cap_n_ = n;
        }
        if (nnz > cap_nnz_) {
            if (Ai_) (void)hipFree(Ai_);
            if (Ax_) (void)hipFree(Ax_);
            Ai_ = nullptr;
            Ax_ = nullptr;
            if (hipMalloc(&Ai_, sizeof(int) * nnz) != hipSuccess || hipMalloc(&Ax_, sizeof(double) * nnz) != hipSuccess) {
                cap_nnz_ = 0;
                return -1;
            }
            cap_nnz_ = nnz;
        }
        return 0;
    }
};

}  // namespace

LinearSolver* new_gpu_dense_lu(int device) { return new GpuDenseLU(device); }

}  // namespace d4host
