// Luttinger-Tisza spectrum of a pyrochlore J1-J2-J3 XXZ model
// (luttinger_tisza.cpp:30-181): for every k of the L0 x L1 x L2 grid the
// (3 n_sl) x (3 n_sl) Hermitian matrix M(k) = sum over directed bonds
// alpha -> beta of J_bond e^{i q.d} in block (alpha, beta), and its lowest
// eigenvalue.
//
// One workgroup per k builds M(k) in LDS (one thread per sublattice block),
// reduces it to a real symmetric tridiagonal matrix by Householder
// reflections (a unitary similarity: the spectrum is unchanged) and finds the
// lowest eigenvalue by Sturm-sequence multisection, 64 shifts per step, one
// per lane of a wave.  Only lambda_min(k) leaves the device.  The eigenvector
// is wanted at k* alone: M(k*) is built again and diagonalised on the host
// by cyclic complex Jacobi rotations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace pyro {
namespace {

constexpr int kLtDim = 48;  // 3 components x 16 sublattices
constexpr int kLtThreads = 256;

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return double2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return double2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

// Blocks of M(k) for threads t = tid, tid + stride, ... < n_sl^2: thread t
// owns block (a, b) = (t / n_sl, t % n_sl) and sums its bonds.  q = 2 pi
// K' / (cell * L) per axis with the centred K' (K < L/2 ? K : K - L).
// Element (3a + i, 3b + j) is written at M[(3a + i) ld + 3b + j].
__device__ void lt_build(const LtBond* __restrict__ bonds, const int32_t* __restrict__ block_begin, int n_sl,
                         int L0, int L1, int L2, int c0, int c1, int c2, long long k, double2* M, int ld,
                         int tid, int stride) {
    const int K2 = int(k % L2), K1 = int((k / L2) % L1), K0 = int(k / (long long)(L1 * L2));
    const double two_pi = 6.283185307179586476925286766559;
    const double q0 = two_pi * double(2 * K0 < L0 ? K0 : K0 - L0) / double(c0 * L0);
    const double q1 = two_pi * double(2 * K1 < L1 ? K1 : K1 - L1) / double(c1 * L1);
    const double q2 = two_pi * double(2 * K2 < L2 ? K2 : K2 - L2) / double(c2 * L2);
    for (int t = tid; t < n_sl * n_sl; t += stride) {
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
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M[size_t(3 * a + i) * ld + 3 * b + j] = double2{re[3 * i + j], im[3 * i + j]};
    }
}

// M(k) of one k into global memory (row-major, N x N), for the host's
// eigenvector at k*.
__global__ __launch_bounds__(kLtThreads) void lt_build_kernel(const LtBond* __restrict__ bonds,
                                                              const int32_t* __restrict__ block_begin, int n_sl,
                                                              int L0, int L1, int L2, int c0, int c1, int c2,
                                                              long long k, double2* __restrict__ M) {
    lt_build(bonds, block_begin, n_sl, L0, L1, L2, c0, c1, c2, k, M, 3 * n_sl, int(threadIdx.x), int(blockDim.x));
}

// lambda_min(k) for k = k_first + blockIdx.x.
//
// Householder step j (j = 0 .. N-3) on the trailing block B = A[j+1:,
// j+1:] (m = N - j - 1 rows) with x = A[j+1:, j]:
//   sigma = |x|, phase = x_0 / |x_0|, alpha = -phase sigma,
//   v = x - alpha e_0 (|v|^2 = 2 sigma (sigma + |x_0|)), beta = 2 / |v|^2,
//   H = I - beta v v^H (H x = alpha e_0),
//   p = beta B v, K = (beta / 2) v^H p, w = p - K v, B <- B - v w^H - w v^H.
// The tridiagonal has the real diagonal d_j = A_jj and the off-diagonals
// alpha_j, whose phases a diagonal unitary similarity removes: the Sturm
// counts need only |alpha_j|^2 = sigma_j^2.
template <int N>
__global__ __launch_bounds__(kLtThreads) void lt_lowest_kernel(const LtBond* __restrict__ bonds,
                                                               const int32_t* __restrict__ block_begin, int L0,
                                                               int L1, int L2, int c0, int c1, int c2,
                                                               long long k_first, double* __restrict__ lambda_min) {
    constexpr int LD = N + 1;  // row stride of A in LDS
    __shared__ double2 A[N * LD];
    __shared__ double2 v[N], p[N];
    __shared__ double d[N], e2[N];
    __shared__ double beta_sh;
    __shared__ double2 K_sh;
    const int tid = int(threadIdx.x);
    const long long k = k_first + blockIdx.x;
    lt_build(bonds, block_begin, N / 3, L0, L1, L2, c0, c1, c2, k, A, LD, tid, kLtThreads);
    __syncthreads();

    for (int j = 0; j < N - 2; ++j) {
        const int m = N - j - 1;
        const double2* x = A + size_t(j + 1) * LD + j;  // x_r = x[r LD]
        if (tid == 0) {
            double s2 = 0.0;
            for (int r = 0; r < m; ++r) s2 += x[r * LD].x * x[r * LD].x + x[r * LD].y * x[r * LD].y;
            d[j] = A[j * LD + j].x;
            e2[j] = s2;
            double beta = 0.0;
            if (s2 > 0.0) {
                const double sig = sqrt(s2);
                const double2 x0 = x[0];
                const double ax0 = sqrt(x0.x * x0.x + x0.y * x0.y);
                const double2 ph = ax0 > 0.0 ? double2{x0.x / ax0, x0.y / ax0} : double2{1.0, 0.0};
                v[0] = double2{x0.x + ph.x * sig, x0.y + ph.y * sig};
                beta = 1.0 / (sig * (sig + ax0));
            }
            beta_sh = beta;
        }
        if (tid >= 1 && tid < m) v[tid] = x[tid * LD];
        __syncthreads();
        const double beta = beta_sh;
        if (beta == 0.0) {  // x = 0: the column is already reduced
            __syncthreads();  // every thread has read beta_sh before step j + 1 writes it
            continue;
        }
        double2* B = A + size_t(j + 1) * LD + (j + 1);
        // p = beta B v: four threads per row, partial sums over kk = part + 4 i
        {
            const int i = tid >> 2, part = tid & 3;
            double2 acc{0.0, 0.0};
            if (i < m)
                for (int kk = part; kk < m; kk += 4) acc = cadd(acc, cmul(B[i * LD + kk], v[kk]));
            acc.x += __shfl_xor(acc.x, 1);
            acc.y += __shfl_xor(acc.y, 1);
            acc.x += __shfl_xor(acc.x, 2);
            acc.y += __shfl_xor(acc.y, 2);
            if (i < m && part == 0) p[i] = double2{beta * acc.x, beta * acc.y};
        }
        __syncthreads();
        if (tid == 0) {
            double2 s{0.0, 0.0};
            for (int i = 0; i < m; ++i) s = cadd(s, cmulc(v[i], p[i]));
            K_sh = double2{0.5 * beta * s.x, 0.5 * beta * s.y};
        }
        __syncthreads();
        const double2 Kc = K_sh;
        if (tid < m) p[tid] = csub(p[tid], cmul(Kc, v[tid]));  // w
        __syncthreads();
        for (int e = tid; e < m * m; e += kLtThreads) {
            const int i = e / m, kk = e - i * m;
            // B_ik -= v_i conj(w_k) + w_i conj(v_k)
            const double2 t1 = cmulc(p[kk], v[i]), t2 = cmulc(v[kk], p[i]);
            double2& b = B[i * LD + kk];
            b = csub(b, cadd(t1, t2));
        }
        __syncthreads();
    }
    if (tid == 0) {
        d[N - 2] = A[(N - 2) * LD + N - 2].x;
        d[N - 1] = A[(N - 1) * LD + N - 1].x;
        const double2 t = A[(N - 1) * LD + N - 2];
        e2[N - 2] = t.x * t.x + t.y * t.y;
        e2[N - 1] = 0.0;
    }
    __syncthreads();
    if (tid >= 64) return;

    // Gershgorin interval of the tridiagonal, then multisection: lane l
    // counts the eigenvalues below x_l = lo + (hi - lo) (l + 1) / 65 (the
    // negative pivots of T - x_l = L D L^T), and the interval becomes
    // [x_{l-1}, x_l] for the first lane whose count is at least one.
    double lo = 1e300, hi = -1e300, emax = 0.0;
    for (int i = tid; i < N; i += 64) {
        const double r = (i > 0 ? sqrt(e2[i - 1]) : 0.0) + sqrt(e2[i]);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
        emax = fmax(emax, e2[i]);
    }
    for (int off = 1; off < 64; off <<= 1) {
        lo = fmin(lo, __shfl_xor(lo, off));
        hi = fmax(hi, __shfl_xor(hi, off));
        emax = fmax(emax, __shfl_xor(emax, off));
    }
    const double pivmin = 1e-290 * fmax(1.0, emax);
    for (int it = 0; it < 14; ++it) {
        const double w = hi - lo;
        const double x = lo + w * double(tid + 1) / 65.0;
        double q = d[0] - x;
        if (fabs(q) < pivmin) q = -pivmin;
        int cnt = q < 0.0;
        for (int i = 1; i < N; ++i) {
            q = d[i] - x - e2[i - 1] / q;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0.0;
        }
        const unsigned long long mask = __ballot(cnt >= 1);
        if (mask == 0ull) {
            lo = lo + w * 64.0 / 65.0;
        } else {
            const int l = __ffsll((long long)mask) - 1;
            const double nlo = l == 0 ? lo : lo + w * double(l) / 65.0;
            hi = lo + w * double(l + 1) / 65.0;
            lo = nlo;
        }
    }
    if (tid == 0) lambda_min[k] = 0.5 * (lo + hi);
}

// Lowest eigenpair of a Hermitian n x n matrix (row-major) by cyclic
// Jacobi rotations.  For the pair (p, q) with b = A_pq = |b| e:
// tau = (A_qq - A_pp) / (2 |b|), t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = 1 / sqrt(1 + t^2), s = t c, and U acts on columns p, q as
//   col_p' = c col_p - s conj(e) col_q,  col_q' = s e col_p + c col_q,
// so that (U^H A U)_pq = 0; V accumulates the rotations.
void hermitian_lowest(std::vector<std::complex<double>> A, int n, double& lam, std::vector<std::complex<double>>& vec) {
    using C = std::complex<double>;
    std::vector<C> V(size_t(n) * n, C(0.0, 0.0));
    for (int i = 0; i < n; ++i) V[size_t(i) * n + i] = 1.0;
    auto a = [&](int r, int c) -> C& { return A[size_t(r) * n + c]; };
    double total = 0.0;
    for (const C& z : A) total += std::norm(z);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += std::norm(a(p, q));
        if (off <= 1e-32 * total) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const C b = a(p, q);
                const double ab = std::abs(b);
                if (ab <= 1e-300) continue;
                const C e = b / ab;
                const double tau = (a(q, q).real() - a(p, p).real()) / (2.0 * ab);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int r = 0; r < n; ++r) {  // A <- A U, V <- V U
                    const C ap = a(r, p), aq = a(r, q);
                    a(r, p) = c * ap - s * std::conj(e) * aq;
                    a(r, q) = s * e * ap + c * aq;
                    const C vp = V[size_t(r) * n + p], vq = V[size_t(r) * n + q];
                    V[size_t(r) * n + p] = c * vp - s * std::conj(e) * vq;
                    V[size_t(r) * n + q] = s * e * vp + c * vq;
                }
                for (int r = 0; r < n; ++r) {  // A <- U^H A
                    const C ap = a(p, r), aq = a(q, r);
                    a(p, r) = c * ap - s * e * aq;
                    a(q, r) = s * std::conj(e) * ap + c * aq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < n; ++i)
        if (a(i, i).real() < a(best, best).real()) best = i;
    lam = a(best, best).real();
    vec.resize(n);
    for (int r = 0; r < n; ++r) vec[r] = V[size_t(r) * n + best];
}

}  // namespace

std::string lt_spectrum(const LtBondTable& table, int n_sl, const int L[3], const int cell[3],
                        double* lambda_min, int32_t* k_star, double* eigvec) {
    if (3 * n_sl != kLtDim) return "the Luttinger-Tisza spectrum supports 16 sublattices (48 x 48 matrices)";
    const std::vector<LtBond>& bonds_by_block = table.bonds;
    const std::vector<int32_t>& block_begin = table.begin;
    const long long n_k = (long long)L[0] * L[1] * L[2];
    if (n_k < 1 || n_k > (1ll << 31) - 1) return "k grid size out of range";
    const int dim = kLtDim;
    hipStream_t stream = nullptr;
    LtBond* d_bonds = nullptr;
    int32_t* d_begin = nullptr;
    double* d_lam = nullptr;
    double2* d_M = nullptr;
    std::string err;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && err.empty()) err = std::string(what) + ": " + hipGetErrorString(e);
        return err.empty();
    };
    double best = 0.0;
    long long k_best = -1;
    do {
        if (!hip_ok(hipStreamCreate(&stream), "hipStreamCreate")) break;
        if (!hip_ok(hipMalloc(&d_bonds, sizeof(LtBond) * std::max<size_t>(1, bonds_by_block.size())), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_begin, sizeof(int32_t) * block_begin.size()), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_lam, sizeof(double) * size_t(n_k)), "hipMalloc")) break;
        if (!hip_ok(hipMalloc(&d_M, sizeof(double2) * dim * dim), "hipMalloc")) break;
        if (!bonds_by_block.empty() &&
            !hip_ok(hipMemcpy(d_bonds, bonds_by_block.data(), sizeof(LtBond) * bonds_by_block.size(),
                              hipMemcpyHostToDevice), "hipMemcpy"))
            break;
        if (!hip_ok(hipMemcpy(d_begin, block_begin.data(), sizeof(int32_t) * block_begin.size(),
                              hipMemcpyHostToDevice), "hipMemcpy"))
            break;
        hipLaunchKernelGGL(lt_lowest_kernel<kLtDim>, dim3(unsigned(n_k)), dim3(kLtThreads), 0, stream, d_bonds,
                           d_begin, L[0], L[1], L[2], cell[0], cell[1], cell[2], 0ll, d_lam);
        if (!hip_ok(hipGetLastError(), "lt_lowest_kernel")) break;
        if (!hip_ok(hipMemcpyAsync(lambda_min, d_lam, sizeof(double) * size_t(n_k), hipMemcpyDeviceToHost, stream),
                    "hipMemcpyAsync"))
            break;
        if (!hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize")) break;
        for (long long k = 0; k < n_k; ++k)
            if (k_best < 0 || lambda_min[k] < best) {
                best = lambda_min[k];
                k_best = k;
            }
        // eigenvector of the lowest eigenvalue at k*
        hipLaunchKernelGGL(lt_build_kernel, dim3(1), dim3(kLtThreads), 0, stream, d_bonds, d_begin, n_sl, L[0], L[1],
                           L[2], cell[0], cell[1], cell[2], k_best, d_M);
        if (!hip_ok(hipGetLastError(), "lt_build_kernel")) break;
        std::vector<double2> Mh(size_t(dim) * dim);
        if (!hip_ok(hipMemcpyAsync(Mh.data(), d_M, sizeof(double2) * dim * dim, hipMemcpyDeviceToHost, stream),
                    "hipMemcpyAsync"))
            break;
        if (!hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize")) break;
        std::vector<std::complex<double>> Mc(Mh.size());
        for (size_t i = 0; i < Mh.size(); ++i) Mc[i] = {Mh[i].x, Mh[i].y};
        double lam = 0.0;
        std::vector<std::complex<double>> vec;
        hermitian_lowest(Mc, dim, lam, vec);
        for (int r = 0; r < dim; ++r) {
            eigvec[2 * r] = vec[r].real();
            eigvec[2 * r + 1] = vec[r].imag();
        }
        k_star[2] = int32_t(k_best % L[2]);
        k_star[1] = int32_t((k_best / L[2]) % L[1]);
        k_star[0] = int32_t(k_best / ((long long)L[1] * L[2]));
    } while (false);
    if (d_M) (void)hipFree(d_M);
    if (d_lam) (void)hipFree(d_lam);
    if (d_begin) (void)hipFree(d_begin);
    if (d_bonds) (void)hipFree(d_bonds);
    if (stream) (void)hipStreamDestroy(stream);
    return err;
}

}  // namespace pyro
