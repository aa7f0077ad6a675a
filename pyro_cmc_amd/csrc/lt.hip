// Luttinger-Tisza spectrum of a pyrochlore J1-J2-J3 XXZ model
// (luttinger_tisza.cpp:30-181): for every k of the L0 x L1 x L2 grid the
// (3 n_sl) x (3 n_sl) Hermitian matrix M(k) = sum over directed bonds
// alpha -> beta of J_bond e^{i q.d} in block (alpha, beta), and its lowest
// eigenvalue.  M(k) is built by a kernel (one workgroup per k, one thread
// per sublattice block) directly in the column-major layout of rocSOLVER's
// batched Hermitian eigensolver, which then returns the spectra.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace pyro {
namespace {

// One workgroup per k; thread t < n_sl^2 owns block (a, b) = (t / n_sl,
// t % n_sl) and sums its bonds.  q = 2 pi K' / (cell * L) per axis with the
// centred K' (K < L/2 ? K : K - L).
__global__ __launch_bounds__(256) void lt_build_kernel(const LtBond* __restrict__ bonds,
                                                       const int32_t* __restrict__ block_begin, int n_sl,
                                                       int L0, int L1, int L2, int c0, int c1, int c2,
                                                       long long k_first, double2* __restrict__ M) {
    const long long k = k_first + blockIdx.x;
    const int K2 = int(k % L2), K1 = int((k / L2) % L1), K0 = int(k / (long long)(L1 * L2));
    const double two_pi = 6.283185307179586476925286766559;
    const double q0 = two_pi * double(2 * K0 < L0 ? K0 : K0 - L0) / double(c0 * L0);
    const double q1 = two_pi * double(2 * K1 < L1 ? K1 : K1 - L1) / double(c1 * L1);
    const double q2 = two_pi * double(2 * K2 < L2 ? K2 : K2 - L2) / double(c2 * L2);
    const int dim = 3 * n_sl;
    double2* Mk = M + size_t(blockIdx.x) * dim * dim;
    for (int t = threadIdx.x; t < n_sl * n_sl; t += blockDim.x) {
        const int a = t / n_sl, b = t % n_sl;
        double re[9] = {}, im[9] = {};
        for (int i = block_begin[t]; i < block_begin[t + 1]; ++i) {
            const LtBond& bd = bonds[i];
            double s, c;
            sincos(q0 * bd.d[0] + q1 * bd.d[1] + q2 * bd.d[2], &s, &c);
            for (int e = 0; e < 9; ++e) {
                re[e] += bd.J[e] * c;
                im[e] += bd.J[e] * s;
            }
        }
        // element (row 3a + i, column 3b + j) at column * dim + row
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Mk[size_t(3 * b + j) * dim + 3 * a + i] = double2{re[3 * i + j], im[3 * i + j]};
    }
}

std::string rocblas_msg(rocblas_status s) { return std::string("rocSOLVER/rocBLAS status ") + std::to_string(int(s)); }

}  // namespace

std::string lt_spectrum(const LtBondTable& table, int n_sl, const int L[3], const int cell[3],
                        double* lambda_min, int32_t* k_star, double* eigvec) {
    const std::vector<LtBond>& bonds_by_block = table.bonds;
    const std::vector<int32_t>& block_begin = table.begin;
    const long long n_k = (long long)L[0] * L[1] * L[2];
    const int dim = 3 * n_sl;
    const size_t mbytes = size_t(dim) * dim * sizeof(double2);
    const long long chunk = std::max<long long>(1, std::min<long long>(n_k, (512ll << 20) / (long long)mbytes));
    hipStream_t stream = nullptr;
    rocblas_handle handle = nullptr;
    LtBond* d_bonds = nullptr;
    int32_t* d_begin = nullptr;
    double2* d_M = nullptr;
    double *d_D = nullptr, *d_E = nullptr;
    rocblas_int* d_info = nullptr;
    std::string err;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && err.empty()) err = std::string(what) + ": " + hipGetErrorString(e);
        return err.empty();
    };
    auto rb_ok = [&](rocblas_status s, const char* what) {
        if (s != rocblas_status_success && err.empty()) err = std::string(what) + ": " + rocblas_msg(s);
        return err.empty();
    };
    std::vector<double> D(size_t(chunk) * dim);
    double best = 0.0;
    long long k_best = -1;
    do {
        if (!hip_ok(hipStreamCreate(&stream), "hipStreamCreate")) break;
        if (!rb_ok(rocblas_create_handle(&handle), "rocblas_create_handle")) break;
        if (!rb_ok(rocblas_set_stream(handle, stream), "rocblas_set_stream")) break;
        if (!hip_ok(hipMalloc(&d_bonds, sizeof(LtBond) * std::max<size_t>(1, bonds_by_block.size())), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_begin, sizeof(int32_t) * block_begin.size()), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_M, mbytes * chunk), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_D, sizeof(double) * dim * chunk), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_E, sizeof(double) * dim * chunk), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_info, sizeof(rocblas_int) * chunk), "hipMalloc")) break;
        if (!bonds_by_block.empty() &&
            !hip_ok(hipMemcpy(d_bonds, bonds_by_block.data(), sizeof(LtBond) * bonds_by_block.size(),
                              hipMemcpyHostToDevice), "hipMemcpy"))
            break;
        if (!hip_ok(hipMemcpy(d_begin, block_begin.data(), sizeof(int32_t) * block_begin.size(),
                              hipMemcpyHostToDevice), "hipMemcpy"))
            break;
        for (long long k0 = 0; k0 < n_k && err.empty(); k0 += chunk) {
            const long long nb = std::min(chunk, n_k - k0);
            hipLaunchKernelGGL(lt_build_kernel, dim3(unsigned(nb)), dim3(256), 0, stream, d_bonds, d_begin, n_sl,
                               L[0], L[1], L[2], cell[0], cell[1], cell[2], k0, d_M);
            if (!hip_ok(hipGetLastError(), "lt_build_kernel")) break;
            if (!rb_ok(rocsolver_zheevd_strided_batched(handle, rocblas_evect_none, rocblas_fill_lower, dim,
                                                       reinterpret_cast<rocblas_double_complex*>(d_M), dim,
                                                       rocblas_stride(dim) * dim, d_D, dim, d_E, dim, d_info,
                                                       rocblas_int(nb)),
                       "rocsolver_zheevd_strided_batched"))
                break;
            if (!hip_ok(hipMemcpyAsync(D.data(), d_D, sizeof(double) * dim * nb, hipMemcpyDeviceToHost, stream),
                        "hipMemcpyAsync"))
                break;
            if (!hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize")) break;
            for (long long i = 0; i < nb; ++i) {
                const double l = D[size_t(i) * dim];  // ascending: the lowest first
                lambda_min[k0 + i] = l;
                if (k_best < 0 || l < best) {
                    best = l;
                    k_best = k0 + i;
                }
            }
        }
        if (!err.empty() || k_best < 0) break;
        // eigenvector of the lowest eigenvalue at k*
        hipLaunchKernelGGL(lt_build_kernel, dim3(1), dim3(256), 0, stream, d_bonds, d_begin, n_sl, L[0], L[1], L[2],
                           cell[0], cell[1], cell[2], k_best, d_M);
        if (!hip_ok(hipGetLastError(), "lt_build_kernel")) break;
        if (!rb_ok(rocsolver_zheevd(handle, rocblas_evect_original, rocblas_fill_lower, dim,
                                    reinterpret_cast<rocblas_double_complex*>(d_M), dim, d_D, d_E, d_info),
                   "rocsolver_zheevd"))
            break;
        std::vector<double2> V(size_t(dim) * dim);
        if (!hip_ok(hipMemcpyAsync(V.data(), d_M, mbytes, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync")) break;
        if (!hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize")) break;
        for (int r = 0; r < dim; ++r) {  // column 0 of the column-major eigenvector matrix
            eigvec[2 * r] = V[r].x;
            eigvec[2 * r + 1] = V[r].y;
        }
        k_star[2] = int32_t(k_best % L[2]);
        k_star[1] = int32_t((k_best / L[2]) % L[1]);
        k_star[0] = int32_t(k_best / ((long long)L[1] * L[2]));
    } while (false);
    if (d_info) (void)hipFree(d_info);
    if (d_E) (void)hipFree(d_E);
    if (d_D) (void)hipFree(d_D);
    if (d_M) (void)hipFree(d_M);
    if (d_begin) (void)hipFree(d_begin);
    if (d_bonds) (void)hipFree(d_bonds);
    if (handle) (void)rocblas_destroy_handle(handle);
    if (stream) (void)hipStreamDestroy(stream);
    return err;
}

}  // namespace pyro
