// HIP kernels for gfx950 (MI355X): graph-coloured single-site sweeps,
// local fields, energy, replica exchange, per-sublattice DFT, contracted
// structure-factor correlators and the sublattice-resolved MFMA products.
//
// Data layout in HBM (per context):
//   spins      S[r][i][c]      Real, c = x,y,z; i = sl*n_cells + cell
//   lifted dir dir[r][i]       int8
//   neighbours nbr[k][i]       int32, slot-major so a wave's loads coalesce
//   couplings  cpl[sl][k]      Real scalar (isotropic) or [sl][k][9]
// A colour is a set of (sublattice, residue) classes with no bond among
// them, so all its sites update concurrently with the same result as any
// sequential order over them (MC.cpp:155-163 visits storage order).

#include <hip/hip_runtime.h>
#include <math.h>

#include "device_update.hpp"
#include "kernels.hpp"
#include "rng.hpp"

namespace pyro {

namespace {

constexpr int kBlock = 256;

__device__ inline int decode_site(int desc, int m, int n_cells, int L1, int L2, int p0,
                                  int p1, int p2) {
    const int sl = desc & 0xFFFF;
    if (p0 == 1 && p1 == 1 && p2 == 1) return sl * n_cells + m;
    const int r0 = (desc >> 16) & 15, r1 = (desc >> 20) & 15, r2 = (desc >> 24) & 15;
    const int q2 = L2 / p2, q1 = L1 / p1;
    const int m2 = m % q2;
    const int t = m / q2;
    const int m1 = t % q1;
    const int m0 = t / q1;
    return sl * n_cells + ((r0 + p0 * m0) * L1 + (r1 + p1 * m1)) * L2 + (r2 + p2 * m2);
}

// Local field h_i = sum_k J_k S_{nbr(k)} of site i in replica r (MC.cpp:87-96).
template <typename Real, int NS, bool ISO>
__device__ inline void local_field(const DeviceState<Real>& st, const Real* __restrict__ Sr, int site, int sl,
                                   Real& hx, Real& hy, Real& hz) {
    unsigned nb[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) nb[k] = unsigned(ld32(st.nbr, unsigned(k * st.N + site)));
    hx = Real(0);
    hy = Real(0);
    hz = Real(0);
    if constexpr (ISO) {
        const Real* __restrict__ c = st.cpl + sl * NS;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const Real ck = c[k];
            const Vec3<Real> v = ld3(Sr, nb[k]);
            hx += ck * v.x;
            hy += ck * v.y;
            hz += ck * v.z;
        }
    } else {
        const Real* __restrict__ J = st.cpl + size_t(sl) * NS * 9;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const Vec3<Real> v = ld3(Sr, nb[k]);
            const Real x = v.x, y = v.y, z = v.z;
            const Real* m = J + 9 * k;
            hx += m[0] * x + m[1] * y + m[2] * z;
            hy += m[3] * x + m[4] * y + m[5] * z;
            hz += m[6] * x + m[7] * y + m[8] * z;
        }
    }
}

template <typename Real, int NS, bool ISO, int KIND>
__device__ inline bool update_site(const DeviceState<Real>& st, const ReplicaParams<Real>& p, int site, int sl,
                                   int r, unsigned long long pass) {
    Real* __restrict__ Sr = st.S + size_t(r) * 3 * size_t(st.N);
    Real hx, hy, hz;
    local_field<Real, NS, ISO>(st, Sr, site, sl, hx, hy, hz);
    return apply_update_p<Real, KIND>(p, st, Sr, site, r, pass, hx, hy, hz);
}

// Strategy 1: one launch per colour; grid.x = classes x blocks-per-class,
// grid.y = replicas.
template <typename Real, int NS, bool ISO, int KIND>
__global__ __launch_bounds__(kBlock) void sweep_colour_kernel(DeviceState<Real> st,
                                                              const int32_t* __restrict__ cls,
                                                              int spc, int bpc,
                                                              unsigned long long pass_index) {
    unsigned bx, ru;
    xcd_split(blockIdx.x, gridDim.x / unsigned(st.R), unsigned(st.R), bx, ru);
    const int ci = int(bx) / bpc;
    const int m = (int(bx) % bpc) * kBlock + threadIdx.x;
    const int r = int(ru);
    const int desc = __builtin_amdgcn_readfirstlane(cls[ci]);
    const int sl = desc & 0xFFFF;
    const unsigned long long pass =
        pass_index + (st.pass_base ? *st.pass_base : 0ull);
    bool accepted = false;
    if (m < spc) {
        const int site = decode_site(desc, m, st.n_cells, st.L1, st.L2, st.p0, st.p1, st.p2);
        accepted = update_site<Real, NS, ISO, KIND>(st, replica_params(st, r), site, sl, r, pass);
    }
    if constexpr (KIND != 3) count_accepts(st, r, bx, accepted);
}

// Strategy 2: one workgroup per replica runs whole sweeps; colours are
// separated by workgroup barriers (the workgroup owns the replica, so no
// other workgroup touches its spins).
template <typename Real, int NS, bool ISO, int FIRST, int BS>
__global__ __launch_bounds__(BS) void sweep_resident_kernel(
    DeviceState<Real> st, const int32_t* __restrict__ all_cls,
    const int32_t* __restrict__ colour_begin, int n_colours, int n_or, int n_sweeps,
    unsigned long long pass_index) {
    const int r = blockIdx.x;
    const int spc = st.n_cells / (st.p0 * st.p1 * st.p2);
    unsigned long long pass = pass_index + (st.pass_base ? *st.pass_base : 0ull);
    const ReplicaParams<Real> rp = replica_params(st, r);
    unsigned int my_acc = 0;
    for (int sw = 0; sw < n_sweeps; ++sw) {
        for (int ps = 0; ps <= n_or; ++ps, ++pass) {
            for (int col = 0; col < n_colours; ++col) {
                const int cb = colour_begin[col], ce = colour_begin[col + 1];
                for (int ci = cb; ci < ce; ++ci) {
                    const int desc = __builtin_amdgcn_readfirstlane(all_cls[ci]);
                    const int sl = desc & 0xFFFF;
                    for (int m = threadIdx.x; m < spc; m += BS) {
                        const int site =
                            decode_site(desc, m, st.n_cells, st.L1, st.L2, st.p0, st.p1, st.p2);
                        if (ps == 0) {
                            my_acc += update_site<Real, NS, ISO, FIRST>(st, rp, site, sl, r, pass) ? 1u : 0u;
                        } else {
                            update_site<Real, NS, ISO, 3>(st, rp, site, sl, r, pass);
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    if (st.acc == nullptr) return;  // uniform: counts not requested
    // block reduction of the accept count
    __shared__ unsigned int red[BS / 64];
    unsigned int v = my_acc;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < BS / 64; ++i) t += red[i];
        atomicAdd(st.acc + r, t);
    }
}

template <typename Real, int NS, bool ISO>
__global__ __launch_bounds__(kBlock) void fields_kernel(DeviceState<Real> st, int r,
                                                        Real* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= st.N) return;
    const size_t N = size_t(st.N);
    const Real* Sr = st.S + size_t(r) * 3 * N;
    const int sl = i / st.n_cells;
    Real hx, hy, hz;
    local_field<Real, NS, ISO>(st, Sr, i, sl, hx, hy, hz);
    out[i] = hx;
    out[N + i] = hy;
    out[2 * N + i] = hz;
}

__device__ inline double block_sum(double v) {
    __shared__ double red[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) t += red[i];
    return t;
}

// Energy terms 1/2 S.h - B.S per site (MC.cpp:305-312), block partial sums
// partial[r][block]; 1-D grid of blocks x R placed XCD-aware (xcd_split).
template <typename Real, int NS, bool ISO>
__global__ __launch_bounds__(kBlock) void energy_partial_kernel(DeviceState<Real> st,
                                                                double* __restrict__ partial) {
    const unsigned gx = gridDim.x / unsigned(st.R);
    unsigned bx, ru;
    xcd_split(blockIdx.x, gx, unsigned(st.R), bx, ru);
    const int i = int(bx) * kBlock + threadIdx.x;
    const int r = int(ru);
    double e = 0.0;
    if (i < st.N) {
        const Real* Sr = st.S + size_t(r) * 3 * size_t(st.N);
        const int sl = i / st.n_cells;
        Real hx, hy, hz;
        local_field<Real, NS, ISO>(st, Sr, i, sl, hx, hy, hz);
        const Vec3<Real> si = ld3(Sr, unsigned(i));
        const Real sx = si.x, sy = si.y, sz = si.z;
        const Real half = Real(0.5) * (sx * hx + sy * hy + sz * hz);
        const Real bs = sx * st.B[3 * r] + sy * st.B[3 * r + 1] + sz * st.B[3 * r + 2];
        e = double(half) - double(bs);
    }
    const double t = block_sum(e);
    if (threadIdx.x == 0) partial[size_t(r) * gx + bx] = t;
}

__global__ __launch_bounds__(kBlock) void energy_finish_kernel(
    const double* __restrict__ partial, int nblk, double n_cells, double* __restrict__ e,
    const int32_t* __restrict__ slot_of, double* Esum, double* E2sum,
    unsigned long long* nE) {
    const int r = blockIdx.x;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) v += partial[size_t(r) * nblk + b];
    const double t = block_sum(v);
    if (threadIdx.x == 0) {
        const double epc = t / n_cells;
        e[r] = epc;
        if (Esum) {
            const int s = slot_of[r];
            Esum[s] += epc;
            E2sum[s] += epc * epc;
            nE[s] += 1ull;
        }
    }
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void randomize_kernel(DeviceState<Real> st, uint32_t k0,
                                                           uint32_t k1) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= st.N) return;
    const Words4 w = philox4x32_10(Words4{uint32_t(i), uint32_t(r), 0u, 0u}, k0, k1);
    const Uniform5 u = split_uniform5(w);
    const double z = 2.0 * u24<double>(u.m[0]) - 1.0;
    const double phi = 6.283185307179586 * u24<double>(u.m[1]);
    const double s = sqrt(1.0 - z * z);
    const size_t N = size_t(st.N);
    st3(st.S + size_t(r) * 3 * N, unsigned(i), Real(s * cos(phi)), Real(s * sin(phi)), Real(z));
    st.dir[size_t(r) * N + i] = 1;
}

// Replica exchange (temper.cpp:437-473): pairs (s, s+1), s = parity mod 2,
// arg = (1/T_s - 1/T_{s+1}) (E_s - E_{s+1}) with total energies; accept if
// arg >= 0 or U < exp(arg).  Pairs are disjoint, so one thread per pair.
__global__ __launch_bounds__(kBlock) void exchange_kernel(ExchangeArgs a) {
    const unsigned long long round = *a.round;
    // exchange x = round / every happens at the end of rounds every - 1,
    // 2 every - 1, ...; its parity and Philox counter come from x (x = round
    // when every = 1)
    const unsigned long long x = round / a.every;
    const bool swap_now = (round + 1ull) % a.every == 0ull;
    const int parity = int(x & 1ull);
    for (int s = parity + 2 * int(threadIdx.x); swap_now && s + 1 < a.R; s += 2 * kBlock) {
        const int ra = a.rep_of[s], rb = a.rep_of[s + 1];
        const double Ea = a.e[ra] * a.n_cells, Eb = a.e[rb] * a.n_cells;
        const double arg = (1.0 / a.T_slot[s] - 1.0 / a.T_slot[s + 1]) * (Ea - Eb);
        const Words4 w = philox4x32_10(Words4{uint32_t(s), 0xFFFFFFFFu, uint32_t(x), uint32_t(x >> 32)}, a.k0, a.k1);
        const bool acc = arg >= 0.0 || u53(w.x, w.y) < exp(arg);
        a.att[s] += 1ull;
        if (acc) {
            a.accp[s] += 1ull;
            a.rep_of[s] = rb;
            a.rep_of[s + 1] = ra;
            a.slot_of[rb] = s;
            a.slot_of[ra] = s + 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *a.round = round + 1ull;
        if (a.pass_base) *a.pass_base += a.passes_per_round;
    }
}

__global__ void bump_kernel(unsigned long long* ctr, unsigned long long n) { *ctr += n; }

// out[r] = map[slot_of[r]] (tempering slot -> accumulator slot, -1 = skip).
__global__ void map_slots_kernel(const int32_t* __restrict__ slot_of, const int32_t* __restrict__ map,
                                 int32_t* __restrict__ out, int R) {
    for (int r = threadIdx.x; r < R; r += blockDim.x) out[r] = map[slot_of[r]];
}

__global__ __launch_bounds__(kBlock) void zero_kernel(uint32_t* p, size_t n) {
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += size_t(gridDim.x) * kBlock)
        p[i] = 0u;
}

__global__ __launch_bounds__(kBlock) void reduce_accepts_kernel(const unsigned int* __restrict__ acc_wave,
                                                                int n, unsigned long long* __restrict__ acc) {
    const int r = blockIdx.x;
    unsigned long long t = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) t += acc_wave[size_t(r) * n + i];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    __shared__ unsigned long long red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long v = 0;
        for (int i = 0; i < kBlock / 64; ++i) v += red[i];
        acc[r] += v;
    }
}

// ---------------------------------------------------------------------------
// Structure factor: separable DFT along one axis of every (r, a, mu) field.
// out[e] = sum_n in[base + n*stride] * W^{K n},  W = exp(-2 pi i / La).
// ---------------------------------------------------------------------------
template <typename Real, bool REAL_IN>
__global__ __launch_bounds__(kBlock) void dft_axis_kernel(const Real* __restrict__ in,
                                                          Real* __restrict__ out,
                                                          const Real* __restrict__ tw, int La,
                                                          int stride, long long total) {
    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (e >= total) return;
    const int K = int((e / stride) % La);
    const long long base = e - (long long)K * stride;
    Real re = Real(0), im = Real(0);
    int j = 0;
    for (int n = 0; n < La; ++n) {
        const Real wr = tw[2 * j], wi = tw[2 * j + 1];
        if constexpr (REAL_IN) {
            const Real x = in[base + (long long)n * stride];
            re += x * wr;
            im += x * wi;
        } else {
            const Real xr = in[2 * (base + (long long)n * stride)];
            const Real xi = in[2 * (base + (long long)n * stride) + 1];
            re += xr * wr - xi * wi;
            im += xr * wi + xi * wr;
        }
        j += K;
        if (j >= La) j -= La;
    }
    out[2 * e] = re;
    out[2 * e + 1] = im;
}

// Line DFT staged through LDS.  The fields are viewed as n_cols columns of
// La points: column c = (o, i) with o = c / s, i = c % s holds the points
// o*La*s + n*s + i, n < La (s = stride of the transformed axis).  A
// workgroup loads 64 columns (coalesced: along i when s > 1, along the
// contiguous lines when s == 1) and the twiddle table into LDS; wave w then
// computes outputs K = w, w + 4, ... of all 64 columns, lane = column, with
// the twiddle index K*n mod La advanced incrementally (exact table values,
// no recurrence), and the results leave through LDS in the load's order.
template <typename Real, bool REAL_IN, int KPT>
__global__ __launch_bounds__(256) void dft_lds_kernel(const Real* __restrict__ in, Real* __restrict__ out,
                                                      const Real* __restrict__ tw, int La, int s,
                                                      unsigned n_cols) {
    extern __shared__ __align__(16) unsigned char dft_smem[];
    Real* xr = reinterpret_cast<Real*>(dft_smem);  // [La][64]
    Real* xi = xr + 64 * La;                       // [La][64]
    Real* wr = xi + 64 * La;                       // [La]
    Real* wi = wr + La;                            // [La]
    const unsigned c0 = blockIdx.x * 64u;  // launcher keeps 2 * La * n_cols < 2^31
    const int P = 64 * La;
    for (int n = threadIdx.x; n < La; n += 256) {
        wr[n] = tw[2 * n];
        wi[n] = tw[2 * n + 1];
    }
    for (int e = threadIdx.x; e < P; e += 256) {
        int j, n;
        unsigned a;
        if (s == 1) {  // contiguous lines: element e of the 64-line block
            j = e / La;
            n = e - j * La;
            a = (c0 + j) * La + n;
        } else {       // strided lines: consecutive lanes take consecutive i
            n = e >> 6;
            j = e & 63;
            const unsigned c = c0 + j, o = c / unsigned(s);
            a = (o * La + n) * unsigned(s) + (c - o * unsigned(s));
        }
        Real re = Real(0), im = Real(0);
        if (c0 + j < n_cols) {
            if constexpr (REAL_IN) {
                re = in[a];
            } else {
                re = in[2 * a];
                im = in[2 * a + 1];
            }
        }
        xr[n * 64 + j] = re;
        xi[n * 64 + j] = im;
    }
    __syncthreads();

    const int j = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Real ar[KPT], ai[KPT];
    int idx[KPT];
#pragma unroll
    for (int t = 0; t < KPT; ++t) {
        ar[t] = Real(0);
        ai[t] = Real(0);
        idx[t] = 0;
    }
    for (int n = 0; n < La; ++n) {
        const Real x = xr[n * 64 + j], y = xi[n * 64 + j];
#pragma unroll
        for (int t = 0; t < KPT; ++t) {
            const int K = w + 4 * t;
            if (K < La) {
                const Real c = wr[idx[t]], d = wi[idx[t]];
                ar[t] += x * c - y * d;
                ai[t] += x * d + y * c;
                idx[t] += K;
                if (idx[t] >= La) idx[t] -= La;
            }
        }
    }
    __syncthreads();  // every read of xr/xi is done: reuse them for the outputs
#pragma unroll
    for (int t = 0; t < KPT; ++t) {
        const int K = w + 4 * t;
        if (K < La) {
            xr[K * 64 + j] = ar[t];
            xi[K * 64 + j] = ai[t];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < P; e += 256) {
        int jj, K;
        unsigned a;
        if (s == 1) {
            jj = e / La;
            K = e - jj * La;
            a = (c0 + jj) * La + K;
        } else {
            K = e >> 6;
            jj = e & 63;
            const unsigned c = c0 + jj, o = c / unsigned(s);
            a = (o * La + K) * unsigned(s) + (c - o * unsigned(s));
        }
        if (c0 + jj < n_cols) {
            out[2 * a] = xr[K * 64 + jj];
            out[2 * a + 1] = xi[K * 64 + jj];
        }
    }
}

typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 line DFT as a complex GEMM on MFMA: for 64 columns of La points
// (same column view as dft_lds_kernel), Y = W X with W[K][n] = e^{-2 pi i K n
// / La}, as real products Yr = Wr Xr - Wi Xi, Yi = Wr Xi + Wi Xr on
// v_mfma_f32_16x16x4f32.  M (= K) is padded to 16 * MT, the reduction (= n)
// to a multiple of 4, both with zeros.  Wave w owns output columns
// 16w .. 16w + 15; X and W^T sit in LDS with bank-conflict-free strides.
template <bool REAL_IN, int MT>
__global__ __launch_bounds__(256) void dft_mfma_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                       const float* __restrict__ tw, int La, int s,
                                                       unsigned n_cols) {
    constexpr int XS = 80;       // X row stride (floats): rows 4 apart land 16 banks apart
    constexpr int P = 16 * MT;   // padded output rows
    extern __shared__ __align__(16) unsigned char dft_smem[];
    const int Kp = (La + 3) & ~3;  // padded reduction length
    float* xr = reinterpret_cast<float*>(dft_smem);  // [Kp][XS]
    float* xi = xr + Kp * XS;                        // [Kp][XS]
    float* wtr = xi + Kp * XS;                       // [Kp][P]  W^T, real part
    float* wti = wtr + Kp * P;                       // [Kp][P]  W^T, imaginary part
    const unsigned c0 = blockIdx.x * 64u;  // launcher keeps 2 * La * n_cols < 2^31
    // Staging in two phases over a fixed 16 rounds (Kp, P <= 64): every
    // round's global load is issued before the first LDS store waits for
    // one (a loop that stored each value as it arrived made the kernel
    // latency-bound).
    float vr[16], vi[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = threadIdx.x + 256 * it;
        vr[it] = 0.f;
        vi[it] = 0.f;
        if (e < Kp * P) {
            const int n = e / P, K = e - n * P;
            if (n < La && K < La) {
                const int j = (K * n) % La;
                vr[it] = tw[2 * j];
                vi[it] = tw[2 * j + 1];
            }
        }
    }
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = threadIdx.x + 256 * it;
        if (e < Kp * P) {
            wtr[e] = vr[it];
            wti[e] = vi[it];
        }
    }
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = threadIdx.x + 256 * it;
        vr[it] = 0.f;
        vi[it] = 0.f;
        if (e < Kp * 64) {
            int j, n;
            unsigned a;
            if (s == 1) {
                j = e / Kp;
                n = e - j * Kp;
                a = (c0 + j) * La + n;
            } else {
                n = e >> 6;
                j = e & 63;
                const unsigned c = c0 + j, o = c / unsigned(s);
                a = (o * La + n) * unsigned(s) + (c - o * unsigned(s));
            }
            if (n < La && c0 + j < n_cols) {
                if constexpr (REAL_IN) {
                    vr[it] = in[a];
                } else {
                    const float2 v = reinterpret_cast<const float2*>(in)[a];  // one 8-byte access
                    vr[it] = v.x;
                    vi[it] = v.y;
                }
            }
        }
    }
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = threadIdx.x + 256 * it;
        if (e < Kp * 64) {
            const int j = s == 1 ? e / Kp : (e & 63);
            const int n = s == 1 ? e - j * Kp : (e >> 6);
            xr[n * XS + j] = vr[it];
            xi[n * XS + j] = vi[it];
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int kq = lane >> 4, ci = lane & 15;
    v4f dr[MT], di[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) dr[m] = di[m] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < Kp; k0 += 4) {
        const int n = k0 + kq;
        const float br = xr[n * XS + 16 * w + ci];
        const float bi = REAL_IN ? 0.f : xi[n * XS + 16 * w + ci];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const float ar = wtr[n * P + 16 * m + ci];
            const float ai = wti[n * P + 16 * m + ci];
            dr[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, dr[m], 0, 0, 0);
            di[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, di[m], 0, 0, 0);
            if constexpr (!REAL_IN) {
                dr[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, dr[m], 0, 0, 0);
                di[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, di[m], 0, 0, 0);
            }
        }
    }
    __syncthreads();  // every read of xr/xi is done: reuse them for the outputs
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int K = 16 * m + 4 * kq + v;
            if (K < Kp) {
                xr[K * XS + 16 * w + ci] = dr[m][v];
                xi[K * XS + 16 * w + ci] = di[m][v];
            }
        }
    __syncthreads();
    for (int e = threadIdx.x; e < La * 64; e += 256) {
        int jj, K;
        unsigned a;
        if (s == 1) {
            jj = e / La;
            K = e - jj * La;
            a = (c0 + jj) * La + K;
        } else {
            K = e >> 6;
            jj = e & 63;
            const unsigned c = c0 + jj, o = c / unsigned(s);
            a = (o * La + K) * unsigned(s) + (c - o * unsigned(s));
        }
        if (c0 + jj < n_cols) reinterpret_cast<float2*>(out)[a] = float2{xr[K * XS + jj], xi[K * XS + jj]};
    }
}

// One Cooley-Tukey step for La = 16 M points (M = 2, 3, 4) on the 16
// columns of a wave: with n = M n1 + n2 and K = k1 + 16 k2,
//   Y_n2[k1] = sum_n1 w16^(k1 n1) x[M n1 + n2]       (M complex GEMMs 16 x 16
//                                                       on v_mfma_f32_16x16x4f32),
//   X[k1 + 16 k2] = sum_n2 wM^(n2 k2) wLa^(n2 k1) Y_n2[k1]   (in registers),
// with wP = e^(-2 pi i / P).  A lane's MFMA accumulators hold Y_n2 of the
// same (k1, column) for every n2, so the twiddles and the M-point butterflies
// need no data exchange; the GEMM work is 1/M of the direct product's.
// load(t, q, n, re, im) gives point n = M (4 t + kq) + q of the lane's
// column ci (t and q are compile-time after unrolling); on return lane
// (kq, ci) holds X[4 kq + v + 16 q] of its column in (yr, yi)[q][v].
// tw[j] = e^(-2 pi i j / La) as (re, im).
template <bool REAL_IN, int M, typename Load>
__device__ __forceinline__ void ct16_lines(Load&& load, const float* __restrict__ tw, int kq, int ci,
                                           float (&yr)[M][4], float (&yi)[M][4]) {
    static_assert(M >= 2 && M <= 4, "lines of 32, 48 or 64 points");
    constexpr int La = 16 * M;
    // A operands: w16^(k1 n1) for row k1 = ci and n1 = 4 t + kq
    float Ar[4], Ai[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int jt = (M * ci * (4 * t + kq)) % La;
        Ar[t] = tw[2 * jt];
        Ai[t] = tw[2 * jt + 1];
    }
    v4f dr[M], di[M];
#pragma unroll
    for (int q = 0; q < M; ++q) dr[q] = di[q] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int q = 0; q < M; ++q) {
            float br, bi;
            load(t, q, M * (4 * t + kq) + q, br, bi);
            dr[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ar[t], br, dr[q], 0, 0, 0);
            di[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ai[t], br, di[q], 0, 0, 0);
            if constexpr (!REAL_IN) {
                dr[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(-Ai[t], bi, dr[q], 0, 0, 0);
                di[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ar[t], bi, di[q], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int k1 = 4 * kq + v;
        float ur[M], ui[M];
        ur[0] = dr[0][v];
        ui[0] = di[0][v];
#pragma unroll
        for (int q = 1; q < M; ++q) {
            const float c = tw[2 * (q * k1)], d = tw[2 * (q * k1) + 1];
            ur[q] = dr[q][v] * c - di[q][v] * d;
            ui[q] = dr[q][v] * d + di[q][v] * c;
        }
        if constexpr (M == 2) {
            yr[0][v] = ur[0] + ur[1];
            yi[0][v] = ui[0] + ui[1];
            yr[1][v] = ur[0] - ur[1];
            yi[1][v] = ui[0] - ui[1];
        } else if constexpr (M == 3) {
            // w3 = -1/2 - i sqrt(3)/2: X1 = u0 + w3 u1 + w3^2 u2, X2 = u0 + w3^2 u1 + w3 u2
            constexpr float h = 0.8660254037844386f;  // sqrt(3) / 2
            const float sr = ur[1] + ur[2], si = ui[1] + ui[2];
            const float dr_ = ur[1] - ur[2], di_ = ui[1] - ui[2];
            yr[0][v] = ur[0] + sr;
            yi[0][v] = ui[0] + si;
            const float mr = ur[0] - 0.5f * sr, mi = ui[0] - 0.5f * si;
            yr[1][v] = mr + h * di_;
            yi[1][v] = mi - h * dr_;
            yr[2][v] = mr - h * di_;
            yi[2][v] = mi + h * dr_;
        } else {
            // w4 = -i
            const float ar = ur[0] + ur[2], ai = ui[0] + ui[2];
            const float br = ur[0] - ur[2], bi = ui[0] - ui[2];
            const float cr = ur[1] + ur[3], cim = ui[1] + ui[3];
            const float er = ur[1] - ur[3], ei = ui[1] - ui[3];
            yr[0][v] = ar + cr;
            yi[0][v] = ai + cim;
            yr[1][v] = br + ei;
            yi[1][v] = bi - er;
            yr[2][v] = ar - cr;
            yi[2][v] = ai - cim;
            yr[3][v] = br - ei;
            yi[3][v] = bi + er;
        }
    }
}

// FP32 line DFT of La = 16 M points (M = 2, 3, 4) by ct16_lines, for the
// same 64-column view as dft_mfma_kernel (wave w owns columns 16 w .. 16 w +
// 15), with no La x La table.
template <bool REAL_IN, int M>
__global__ __launch_bounds__(256) void dft_ct_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                     const float* __restrict__ tw, int s, unsigned n_cols) {
    constexpr int La = 16 * M;
    // row stride (floats): the B-operand rows M apart (kq = 0..3) start 16
    // banks apart, so a wave's reads hit 64 distinct banks
    constexpr int XS = M == 2 ? 72 : (M == 3 ? 80 : 68);
    constexpr int ROUNDS = La * 64 / 256;  // staging: 4 M elements per thread
    __shared__ float xr[La * XS];
    __shared__ float xi[La * XS];
    const unsigned c0 = blockIdx.x * 64u;  // launcher keeps 2 * La * n_cols < 2^31
    auto where = [&](int e, int& j, int& n, unsigned& a) __attribute__((always_inline)) {
        if (s == 1) {
            j = e / La;
            n = e - j * La;
            a = (c0 + j) * La + n;
        } else {
            n = e >> 6;
            j = e & 63;
            const unsigned c = c0 + j, o = c / unsigned(s);
            a = (o * La + n) * unsigned(s) + (c - o * unsigned(s));
        }
    };
    // every global load issued before the first LDS store
    float vr[ROUNDS], vi[ROUNDS];
#pragma unroll
    for (int it = 0; it < ROUNDS; ++it) {
        int j, n;
        unsigned a;
        where(threadIdx.x + 256 * it, j, n, a);
        vr[it] = 0.f;
        vi[it] = 0.f;
        if (c0 + j < n_cols) {
            if constexpr (REAL_IN) {
                vr[it] = in[a];
            } else {
                const float2 v = reinterpret_cast<const float2*>(in)[a];
                vr[it] = v.x;
                vi[it] = v.y;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < ROUNDS; ++it) {
        int j, n;
        unsigned a;
        where(threadIdx.x + 256 * it, j, n, a);
        xr[n * XS + j] = vr[it];
        xi[n * XS + j] = vi[it];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int kq = lane >> 4, ci = lane & 15;
    float yr[M][4], yi[M][4];
    ct16_lines<REAL_IN, M>(
        [&](int, int, int n, float& br, float& bi) __attribute__((always_inline)) {
            br = xr[n * XS + 16 * w + ci];
            bi = REAL_IN ? 0.f : xi[n * XS + 16 * w + ci];
        },
        tw, kq, ci, yr, yi);
    __syncthreads();  // every B read is done: reuse xr/xi for the outputs
#pragma unroll
    for (int q = 0; q < M; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int K = 4 * kq + v + 16 * q;
            xr[K * XS + 16 * w + ci] = yr[q][v];
            xi[K * XS + 16 * w + ci] = yi[q][v];
        }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ROUNDS; ++it) {
        int jj, K;
        unsigned a;
        where(threadIdx.x + 256 * it, jj, K, a);
        if (c0 + jj < n_cols) reinterpret_cast<float2*>(out)[a] = float2{xr[K * XS + jj], xi[K * XS + jj]};
    }
}

// Complex lines of La = 16 M points with stride s > 1 (axis 0, or axis 1
// of a lattice the plane kernel does not take): each lane loads its MFMA
// B operands straight from global memory and stores its outputs straight
// back, no LDS and no barrier.  For each (t, q) a wave reads rows n = M (4 t
// + kq) + q of its 16 consecutive columns: four 128-byte segments when s is
// a multiple of 16.  Same grid as dft_ct_kernel (wave w of workgroup b owns
// columns 64 b + 16 w .. + 15).
template <int M>
__global__ __launch_bounds__(256) void dft_ct_strided_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                             const float* __restrict__ tw, int s, unsigned n_cols) {
    constexpr int La = 16 * M;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int kq = lane >> 4, ci = lane & 15;
    const unsigned c = blockIdx.x * 64u + 16u * unsigned(w) + unsigned(ci);
    const bool ok = c < n_cols;
    const unsigned o = c / unsigned(s);
    const unsigned base = o * unsigned(La) * unsigned(s) + (c - o * unsigned(s));  // point n at base + n s
    const float2* __restrict__ in2 = reinterpret_cast<const float2*>(in);
    float2 b[4][M];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < M; ++q)
            b[t][q] = ok ? in2[base + unsigned(M * (4 * t + kq) + q) * unsigned(s)] : float2{0.f, 0.f};
    float yr[M][4], yi[M][4];
    ct16_lines<false, M>(
        [&](int t, int q, int, float& br, float& bi) __attribute__((always_inline)) {
            br = b[t][q].x;
            bi = b[t][q].y;
        },
        tw, kq, ci, yr, yi);
    if (!ok) return;
    float2* __restrict__ out2 = reinterpret_cast<float2*>(out);
#pragma unroll
    for (int q = 0; q < M; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) out2[base + unsigned(4 * kq + v + 16 * q) * unsigned(s)] = float2{yr[q][v], yi[q][v]};
}

// FP32 DFT along axes 2 and 1 of one (field, n0) plane of real values, L1 =
// 16 M1 by L2 = 16 M2 points, in LDS: read once, written once as complex
// (the axis-0 pass follows), so the two passes cost 12 B per point of HBM
// traffic instead of 4 + 8 + 8 + 8.  Pass A transforms the rows (column i1 of
// the Cooley-Tukey step = wave i1 / 16) and stores its result transposed,
// P'[k2][i1]; pass B then reads columns of P' with the same transposed
// access.  The row stride 68 puts a wave's transposed reads on 64 distinct
// banks for 48-point lines (4 ci + 3 kq mod 64 distinct).
//
// The input is real, so only the kept half k2 < H = L2 / 2 + 1 is carried
// into pass B and written (plane stride L1 H): the waves of pass B take the
// columns k2 < 16 ceil(H / 16).
template <int M1, int M2>
__global__ __launch_bounds__(256) void dft_plane_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                        const float* __restrict__ tw1,
                                                        const float* __restrict__ tw2, int n_sl, int L0) {
    constexpr int L1 = 16 * M1, L2 = 16 * M2, NP = L1 * L2;
    constexpr int H = L2 / 2 + 1, NPH = L1 * H, WB = (H + 15) / 16;
    constexpr int XS = (M1 == 3 && M2 == 3) ? 68 : 71;
    constexpr int ROWS = L1 > L2 ? L1 : L2;
    constexpr int RND = (NP + 255) / 256;
    __shared__ float pr[ROWS * XS];
    __shared__ float pi[ROWS * XS];
    const size_t base = size_t(blockIdx.x) * NPH;  // plane p = field * L0 + n0 of the output
    // input: field = (r * 3 + a) * n_sl + mu of spins S[r][i][a], every third value
    const int field = int(blockIdx.x) / L0, n0 = int(blockIdx.x) - field * L0;
    const int mu = field % n_sl, ra = field / n_sl, a = ra % 3, r = ra / 3;
    const size_t nc = size_t(L0) * NP;
    const float* src = in + (size_t(r) * n_sl + mu) * 3 * nc + 3 * size_t(n0) * NP + a;
    float v[RND];
#pragma unroll
    for (int it = 0; it < RND; ++it) {
        const int e = threadIdx.x + 256 * it;
        v[it] = e < NP ? src[3 * e] : 0.f;
    }
#pragma unroll
    for (int it = 0; it < RND; ++it) {
        const int e = threadIdx.x + 256 * it;
        if (e < NP) pr[(e / L2) * XS + e % L2] = v[it];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int kq = lane >> 4, ci = lane & 15;
    // pass A: rows i1 = 16 w + ci (M1 waves), points along n2
    float yr[M2][4], yi[M2][4];
    if (w < M1)
        ct16_lines<true, M2>(
            [&](int, int, int n, float& br, float& bi) __attribute__((always_inline)) {
                br = pr[(16 * w + ci) * XS + n];
                bi = 0.f;
            },
            tw2, kq, ci, yr, yi);
    __syncthreads();  // every read of the real plane is done
    if (w < M1) {
#pragma unroll
        for (int q = 0; q < M2; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k2 = 4 * kq + u + 16 * q;
                if (16 * q < 16 * WB) {  // rows read by pass B
                    pr[k2 * XS + 16 * w + ci] = yr[q][u];
                    pi[k2 * XS + 16 * w + ci] = yi[q][u];
                }
            }
    }
    __syncthreads();
    // pass B: columns k2 = 16 w + ci (WB waves), points along n1
    if (w < WB) {
        float zr[M1][4], zi[M1][4];
        ct16_lines<false, M1>(
            [&](int, int, int n, float& br, float& bi) __attribute__((always_inline)) {
                br = pr[(16 * w + ci) * XS + n];
                bi = pi[(16 * w + ci) * XS + n];
            },
            tw1, kq, ci, zr, zi);
        float2* o = reinterpret_cast<float2*>(out) + base;
        const int k2 = 16 * w + ci;
        if (k2 < H) {
#pragma unroll
            for (int q = 0; q < M1; ++q)
#pragma unroll
                for (int u = 0; u < 4; ++u) o[(4 * kq + u + 16 * q) * H + k2] = float2{zr[q][u], zi[q][u]};
        }
    }
}

// Contracted correlators (ssf_manager.hpp:260-293 with the separable phase
// fold): b^a(K) = sum_mu e^{-i q.r_mu} A^a_mu(K); S_c = conj(b^alpha) b^beta.
constexpr int kContractBlock = 64;  // one wave per workgroup

// full index K of kept index kh (SsfState::H)
template <typename Real>
__device__ __forceinline__ int full_k(const SsfState<Real>& st, int kh) {
    const int l = kh / st.H;
    return l * st.L2 + (kh - l * st.H);
}

// -K of a full index K
template <typename Real>
__device__ __forceinline__ int neg_full_k(const SsfState<Real>& st, int K) {
    const int l = K / st.L2, k2 = K - l * st.L2;
    const int k0 = l / st.L1, k1 = l - k0 * st.L1;
    return ((k0 ? st.L0 - k0 : 0) * st.L1 + (k1 ? st.L1 - k1 : 0)) * st.L2 + (k2 ? st.L2 - k2 : 0);
}

// Whether the partner -K of kept index kh lies outside the kept half with a
// Nyquist index on axis 0 or 1 (k0 = L0 / 2 or k1 = L1 / 2): there the
// centred q of -K is not minus that of K, so S(-K) is not conj(S(K)) and is
// accumulated directly, from A(-K) = conj(A(K)) folded with the phases of
// -K.  The other K outside the kept half are filled in as conj(S(-K)) when
// read.
template <typename Real>
__device__ __forceinline__ bool mirror_direct(const SsfState<Real>& st, int kh) {
    const int l = kh / st.H, k2 = kh - l * st.H;
    if (k2 < 1 || k2 > st.L2 - st.H) return false;
    const int k0 = l / st.L1, k1 = l - k0 * st.L1;
    return 2 * k0 == st.L0 || 2 * k1 == st.L1;
}

// S_c(K) += conj(b^al) b^be (and the squares of its parts) for the
// correlators of one (slot, K).
template <typename Real>
__device__ __forceinline__ void add_correlators(const CorrArgs& c, int slot, size_t nc, int K, const Real (&br)[3],
                                                const Real (&bi)[3]) {
    for (int cc = 0; cc < c.n_corr; ++cc) {
        const int al = c.alpha[cc], be = c.beta[cc];
        // explicit fma: every kernel rounds the product alike
        const double re = double(fma(br[al], br[be], bi[al] * bi[be]));
        const double im = double(fma(br[al], bi[be], -(bi[al] * br[be])));
        double* A = c.acc + 2 * ((size_t(slot) * c.n_corr + cc) * nc + K);
        A[0] += re;
        A[1] += im;
        if (c.acc2) {
            double* A2 = c.acc2 + 2 * ((size_t(slot) * c.n_corr + cc) * nc + K);
            A2[0] += re * re;
            A2[1] += im * im;
        }
    }
}

// Sublattices whose values contract_kernel loads together (the pyrochlore
// cell's 16); lattices with more take the rest one by one.
constexpr int kMaxSl = 16;

template <typename Real>
__global__ __launch_bounds__(kContractBlock) void contract_kernel(SsfState<Real> st, CorrArgs c,
                                                                  int r_begin) {
    const int kh = blockIdx.x * kContractBlock + threadIdx.x;
    const int rl = blockIdx.y;
    const int r = r_begin + rl;
    const int slot = c.slot_of[r];
    if (slot < 0) return;
    if (kh < st.n_kh) {
        const int K = full_k(st, kh);
        Real br[3], bi[3];
        const size_t nc = size_t(st.n_cells), nh = size_t(st.n_kh);
        // the first kMaxSl sublattices' phase and field values loaded
        // together (those past n_sl read sublattice 0's and add nothing), any
        // further sublattices one by one after them; the sums in mu order
        const int ns = st.n_sl;
        auto fold = [&](int Kp, Real sign_i) {
            Real pr[kMaxSl], pi[kMaxSl];
#pragma unroll
            for (int mu = 0; mu < kMaxSl; ++mu) {
                const size_t at = 2 * (size_t(mu < ns ? mu : 0) * nc + size_t(Kp));
                pr[mu] = st.phase[at];
                pi[mu] = st.phase[at + 1];
            }
            for (int a = 0; a < 3; ++a) {
                const Real* F = st.F + 2 * ((size_t(rl) * 3 + a) * ns) * nh;
                Real fr[kMaxSl], fi[kMaxSl];
#pragma unroll
                for (int mu = 0; mu < kMaxSl; ++mu) {
                    const size_t at = 2 * (size_t(mu < ns ? mu : 0) * nh + size_t(kh));
                    fr[mu] = F[at];
                    fi[mu] = F[at + 1];
                }
                Real sr = Real(0), si = Real(0);
#pragma unroll
                for (int mu = 0; mu < kMaxSl; ++mu) {
                    if (mu < ns) {
                        const Real f_r = fr[mu], f_i = sign_i * fi[mu];
                        sr += f_r * pr[mu] - f_i * pi[mu];
                        si += f_r * pi[mu] + f_i * pr[mu];
                    }
                }
                for (int mu = kMaxSl; mu < ns; ++mu) {
                    const Real f_r = F[2 * (mu * nh + kh)], f_i = sign_i * F[2 * (mu * nh + kh) + 1];
                    const Real p_r = st.phase[2 * (mu * nc + Kp)], p_i = st.phase[2 * (mu * nc + Kp) + 1];
                    sr += f_r * p_r - f_i * p_i;
                    si += f_r * p_i + f_i * p_r;
                }
                br[a] = sr;
                bi[a] = si;
            }
        };
        fold(K, Real(1));
        add_correlators(c, slot, nc, K, br, bi);
        if (mirror_direct(st, kh)) {  // A(-K) = conj(A(K)) with the phases of -K
            const int Kn = neg_full_k(st, K);
            fold(Kn, Real(-1));
            add_correlators(c, slot, nc, Kn, br, bi);
        }
    }
    if (kh == 0) c.n_samples[slot] += 1ull;
}


// ---------------------------------------------------------------------------
// Sublattice-resolved correlators C^{ab}_{mu nu}(K) = sum conj(A^a_mu) A^b_nu,
// i.e. the Hermitian 48x48 matrix M = sum_samples a^H a per K.  With
// X = [Re a | Im a] (one row per sample), X^T X holds RR, RI, IR, II and
// Re M = RR + II, Im M = RI - IR.  X^T X runs on FP32 MFMA
// (v_mfma_f32_16x16x4_f32) over the buffered samples and is added to FP64.
// ---------------------------------------------------------------------------

// Stage the transformed fields of replica (r_begin + blockIdx.y) into row
// cnt[slot] of its slot: buf[slot][b][k][2*16*T] = [Re A_f | Im A_f] for the
// features f = a * n_sl + mu (zero-padded to 16*T).  A workgroup takes 64
// consecutive k: reads along k (coalesced), transposes through LDS, and
// writes the 64 rows as one contiguous block.
template <typename Real, int T>
__global__ __launch_bounds__(kBlock) void munu_stage_kernel(SsfState<Real> st,
                                                            const int32_t* __restrict__ slot_of,
                                                            int r_begin, float* __restrict__ buf,
                                                            const unsigned int* __restrict__ cnt,
                                                            int Bmax) {
    constexpr int FW = 2 * 16 * T;
    __shared__ float xs[FW][65];  // [feature][k], padded against bank conflicts
    const int rl = blockIdx.y;
    const int slot = slot_of[r_begin + rl];
    if (slot < 0) return;
    const unsigned int b = cnt[slot];
    if (b >= unsigned(Bmax)) return;
    const int nc = st.n_kh;  // rows of the kept K
    const int K0 = int(blockIdx.x) * 64;
    const int nf = 3 * st.n_sl;
    const int kl = threadIdx.x & 63;
    {
        // a wave's fields f = w, w + 4, ...: every load issued before any
        // value is stored, so that they are in flight together
        constexpr int NFW = 16 * T / (kBlock / 64);
        static_assert(16 * T % (kBlock / 64) == 0, "whole fields per wave");
        // the values kept in their own precision until stored, so that no
        // conversion waits on its load ahead of the next load
        Real re[NFW], im[NFW];
        const int K = K0 + kl;
#pragma unroll
        for (int i = 0; i < NFW; ++i) {
            const int f = int(threadIdx.x >> 6) + i * (kBlock / 64);
            re[i] = Real(0);
            im[i] = Real(0);
            if (f < nf && K < nc) {
                const int a = f / st.n_sl, mu = f % st.n_sl;
                const Real* F = st.F + 2 * (((size_t(rl) * 3 + a) * st.n_sl + mu) * size_t(nc) + K);
                re[i] = F[0];
                im[i] = F[1];
            }
        }
#pragma unroll
        for (int i = 0; i < NFW; ++i) {
            const int f = int(threadIdx.x >> 6) + i * (kBlock / 64);
            xs[f][kl] = float(re[i]);
            xs[16 * T + f][kl] = float(im[i]);
        }
    }
    __syncthreads();
    const int nk = min(64, nc - K0);
    float* dst = buf + ((size_t(slot) * Bmax + b) * size_t(nc) + K0) * FW;
    for (int e = threadIdx.x; e < nk * FW; e += kBlock) {
        const int k = e / FW, f = e - k * FW;
        dst[e] = xs[f][k];
    }
}

// contract_kernel for FP32 contexts, and with STAGE also munu_stage_kernel,
// from one read of the transformed fields: the 64 k of a workgroup are
// transposed through LDS as for the staging (and with STAGE the row block is
// written to the buffer); wave a < 3 then folds component a of the same 64 k
// from LDS (b^a(K) = sum_mu e^{-i q.r_mu} A^a_mu(K), in the order of
// contract_kernel) and every wave updates the correlators cc = w, w + 4, ...
// of its slot.
//
// With cbuf (STAGE only) the products of the kept K go to the buffer as
// FP32 pairs, cbuf[slot][b][c][kh] with b the row the staging writes, and
// corr_flush_kernel adds them to the FP64 sums in sample order when the
// rows are flushed: the same additions as the direct update, without an
// FP64 read-modify-write per sample.  The directly accumulated partners -K
// of the Nyquist planes are updated in place.
template <int T, bool STAGE>
__global__ __launch_bounds__(kBlock) void contract_stage_kernel(SsfState<float> st, CorrArgs c, int r_begin,
                                                                float* __restrict__ buf,
                                                                const unsigned int* __restrict__ cnt, int Bmax,
                                                                float2* __restrict__ cbuf) {
    constexpr int FW = 2 * 16 * T;
    __shared__ float xs[FW][65];  // [feature][k], padded against bank conflicts
    const int rl = blockIdx.y;
    const int slot = c.slot_of[r_begin + rl];
    if (slot < 0) return;
    const int nc = st.n_kh;  // kept K (fields and staged rows); accumulators by full K
    const int K0 = int(blockIdx.x) * 64;
    const int ns = st.n_sl, nf = 3 * ns;
    const int kl = threadIdx.x & 63;
    {
        // a wave's fields f = w, w + 4, ...: every load issued before any
        // value is stored, so that they are in flight together
        constexpr int NFW = 16 * T / (kBlock / 64);
        static_assert(16 * T % (kBlock / 64) == 0, "whole fields per wave");
        float re[NFW], im[NFW];
        const int K = K0 + kl;
#pragma unroll
        for (int i = 0; i < NFW; ++i) {
            const int f = int(threadIdx.x >> 6) + i * (kBlock / 64);
            re[i] = 0.f;
            im[i] = 0.f;
            if (f < nf && K < nc) {
                const int a = f / ns, mu = f % ns;
                const float2 v = *reinterpret_cast<const float2*>(
                    st.F + 2 * (((size_t(rl) * 3 + a) * ns + mu) * size_t(nc) + K));
                re[i] = v.x;
                im[i] = v.y;
            }
        }
#pragma unroll
        for (int i = 0; i < NFW; ++i) {
            const int f = int(threadIdx.x >> 6) + i * (kBlock / 64);
            xs[f][kl] = re[i];
            xs[16 * T + f][kl] = im[i];
        }
    }
    __syncthreads();
    const unsigned int b = STAGE ? cnt[slot] : 0u;
    if (STAGE && b < unsigned(Bmax)) {
        const int nk = min(64, nc - K0);
        float* dst = buf + ((size_t(slot) * Bmax + b) * size_t(nc) + K0) * FW;
        for (int e = threadIdx.x; e < nk * FW; e += kBlock) {
            const int k = e / FW, f = e - k * FW;
            dst[e] = xs[f][k];
        }
    }
    const int kh = K0 + kl;
    const int K = kh < nc ? full_k(st, kh) : 0;
    const bool mir = kh < nc && mirror_direct(st, kh);
    const int Kn = mir ? neg_full_k(st, K) : 0;
    const size_t nfull = size_t(st.n_cells);
    const int w = int(threadIdx.x >> 6);
    __shared__ float bsh[2][2][3][64];  // [K, -K][re, im][component][k]
    if (w < 3 && kh < nc) {
        // the phases of every sublattice loaded together (sublattices past
        // n_sl read sublattice 0's and add nothing), the sums in mu order
        constexpr int MAXNS = 16 * T / 3;  // n_sl <= 16 T / 3 (T = ceil(3 n_sl / 16))
        float2 ph[MAXNS];
#pragma unroll
        for (int mu = 0; mu < MAXNS; ++mu)
            ph[mu] = *reinterpret_cast<const float2*>(st.phase + 2 * ((mu < ns ? mu : 0) * nfull + K));
        float sr = 0.f, si = 0.f;
#pragma unroll
        for (int mu = 0; mu < MAXNS; ++mu) {
            if (mu < ns) {
                const float fr = xs[w * ns + mu][kl], fi = xs[16 * T + w * ns + mu][kl];
                const float pr = ph[mu].x, pi = ph[mu].y;
                sr += fr * pr - fi * pi;
                si += fr * pi + fi * pr;
            }
        }
        bsh[0][0][w][kl] = sr;
        bsh[0][1][w][kl] = si;
        if (mir) {  // A(-K) = conj(A(K)) with the phases of -K
#pragma unroll
            for (int mu = 0; mu < MAXNS; ++mu)
                ph[mu] = *reinterpret_cast<const float2*>(st.phase + 2 * ((mu < ns ? mu : 0) * nfull + Kn));
            sr = 0.f;
            si = 0.f;
#pragma unroll
            for (int mu = 0; mu < MAXNS; ++mu) {
                if (mu < ns) {
                    const float fr = xs[w * ns + mu][kl], fi = -xs[16 * T + w * ns + mu][kl];
                    const float pr = ph[mu].x, pi = ph[mu].y;
                    sr += fr * pr - fi * pi;
                    si += fr * pi + fi * pr;
                }
            }
            bsh[1][0][w][kl] = sr;
            bsh[1][1][w][kl] = si;
        }
    }
    __syncthreads();
    const bool buffered = STAGE && cbuf != nullptr && b < unsigned(Bmax);
    for (int m = 0; m < (mir ? 2 : 1); ++m) {
        if (kh >= nc) break;
        const int Kw = m ? Kn : K;
        const float br[3] = {bsh[m][0][0][kl], bsh[m][0][1][kl], bsh[m][0][2][kl]};
        const float bi[3] = {bsh[m][1][0][kl], bsh[m][1][1][kl], bsh[m][1][2][kl]};
        for (int cc = w; cc < c.n_corr; cc += kBlock / 64) {
            const int al = c.alpha[cc], be = c.beta[cc];
            const float rf = fmaf(br[al], br[be], bi[al] * bi[be]);
            const float imf = fmaf(br[al], bi[be], -(bi[al] * br[be]));
            if (buffered && m == 0) {
                cbuf[((size_t(slot) * Bmax + b) * c.n_corr + cc) * size_t(nc) + kh] = float2{rf, imf};
                continue;
            }
            const double re = double(rf);
            const double im = double(imf);
            double* A = c.acc + 2 * ((size_t(slot) * c.n_corr + cc) * nfull + Kw);
            A[0] += re;
            A[1] += im;
            if (c.acc2) {
                double* A2 = c.acc2 + 2 * ((size_t(slot) * c.n_corr + cc) * nfull + Kw);
                A2[0] += re * re;
                A2[1] += im * im;
            }
        }
    }
    if (K0 == 0 && threadIdx.x == 0) c.n_samples[slot] += 1ull;
}

// The buffered contracted samples b < cnt[slot] of every (slot, correlator,
// kept K), added to the FP64 sums in sample order (the additions the direct
// update makes, sample by sample).  One thread per (correlator, kept K).
__global__ __launch_bounds__(kBlock) void corr_flush_kernel(const float2* __restrict__ cbuf,
                                                            const unsigned int* __restrict__ cnt,
                                                            double* __restrict__ acc, double* __restrict__ acc2,
                                                            int n_corr, int n_kh, int H, int L2, int n_k, int Bmax) {
    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    const int slot = int(blockIdx.y);
    if (e >= (long long)n_corr * n_kh) return;
    const int nb = int(cnt[slot]);
    if (nb == 0) return;
    const int cc = int(e / n_kh), kh = int(e - (long long)cc * n_kh);
    const int l = kh / H, K = l * L2 + (kh - l * H);
    const size_t at = 2 * ((size_t(slot) * n_corr + cc) * size_t(n_k) + K);
    double a0 = acc[at], a1 = acc[at + 1];
    double s0 = 0.0, s1 = 0.0;
    if (acc2) {
        s0 = acc2[at];
        s1 = acc2[at + 1];
    }
    const float2* v = cbuf + (size_t(slot) * Bmax * n_corr + cc) * size_t(n_kh) + kh;
    const size_t step = size_t(n_corr) * n_kh;
    for (int b = 0; b < nb; ++b) {
        const float2 x = v[b * step];
        const double re = double(x.x), im = double(x.y);
        a0 += re;
        a1 += im;
        s0 += re * re;
        s1 += im * im;
    }
    acc[at] = a0;
    acc[at + 1] = a1;
    if (acc2) {
        acc2[at] = s0;
        acc2[at + 1] = s1;
    }
}

__global__ void munu_count_kernel(const int32_t* __restrict__ slot_of, int r_begin, int r_count,
                                  unsigned int* cnt, int Bmax) {
    for (int i = threadIdx.x; i < r_count; i += blockDim.x) {
        const int s = slot_of[r_begin + i];
        if (s >= 0 && cnt[s] < unsigned(Bmax)) atomicAdd(cnt + s, 1u);
    }
}

// One wave per (slot, K): the Hermitian 16T x 16T matrix
// M_fg += sum_b conj(A_f) A_g over the cnt[slot] buffered rows, as real
// products of the rows X = [Re | Im]: Re M = RR + II, Im M = RI - IR.  Only
// the 16-feature tile pairs (I, J) with I <= J are formed and accumulated;
// the lower tiles follow by conjugation when read.
template <int T>
__global__ __launch_bounds__(256) void munu_flush_kernel(const float* __restrict__ buf,
                                                         const unsigned int* __restrict__ cnt,
                                                         double* __restrict__ acc, int n_k,
                                                         int Bmax, int n_f) {
    constexpr int FW = 32 * T;
    constexpr int NP = T * (T + 1) / 2;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int K = blockIdx.x * 4 + wave;
    const int slot = blockIdx.y;
    if (K >= n_k) return;
    const int nb = int(cnt[slot]);
    if (nb == 0) return;
    v4f RR[NP], II[NP], RI[NP], IR[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) RR[p] = II[p] = RI[p] = IR[p] = v4f{0.f, 0.f, 0.f, 0.f};
    const int kk = lane >> 4;   // which of the 4 rows of this step
    const int col = lane & 15;  // feature within a tile
    const size_t row_stride = size_t(n_k) * FW;  // from row b to row b + 1
    const float* X = buf + (size_t(slot) * Bmax * n_k + K) * FW + col;
    // rows of step b0 (4 rows, row b0 + kk on this lane), zero past nb
    auto load_step = [&](int b0, float (&r)[T], float (&i)[T]) __attribute__((always_inline)) {
        const int b = b0 + kk;
#pragma unroll
        for (int I = 0; I < T; ++I) {
            r[I] = (b < nb) ? X[size_t(b) * row_stride + 16 * I] : 0.f;
            i[I] = (b < nb) ? X[size_t(b) * row_stride + 16 * (T + I)] : 0.f;
        }
    };
    // loads run four steps ahead of the MFMAs that consume them (70.5 ->
    // 54.4 ms per flush at L=48 x 16, B = 192, against loads waited on in
    // the same step; two steps ahead: 56.3 ms)
    float xr[T], xi[T], yr[T], yi[T], ur[T], ui[T], wr[T], wi[T];
    load_step(0, xr, xi);
    load_step(4, yr, yi);
    load_step(8, ur, ui);
    load_step(12, wr, wi);
    for (int b0 = 0; b0 < nb; b0 += 4) {
        float zr[T], zi[T];
        load_step(b0 + 16, zr, zi);
        int p = 0;
#pragma unroll
        for (int I = 0; I < T; ++I)
#pragma unroll
            for (int J = I; J < T; ++J, ++p) {
                RR[p] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[I], xr[J], RR[p], 0, 0, 0);
                II[p] = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[I], xi[J], II[p], 0, 0, 0);
                RI[p] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[I], xi[J], RI[p], 0, 0, 0);
                IR[p] = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[I], xr[J], IR[p], 0, 0, 0);
            }
#pragma unroll
        for (int I = 0; I < T; ++I) {
            xr[I] = yr[I];
            xi[I] = yi[I];
            yr[I] = ur[I];
            yi[I] = ui[I];
            ur[I] = wr[I];
            ui[I] = wi[I];
            wr[I] = zr[I];
            wi[I] = zi[I];
        }
    }
    // element v of a tile accumulator on this lane: row 16I + 4*(lane>>4) + v,
    // column 16J + (lane&15); the stored tile pairs (I <= J) are packed,
    // pair p at [p][row % 16][column % 16] (munu_acc_pairs)
    double2* out = reinterpret_cast<double2*>(acc) + (size_t(slot) * n_k + K) * size_t(NP) * 256;
    int p = 0;
#pragma unroll
    for (int I = 0; I < T; ++I)
#pragma unroll
        for (int J = I; J < T; ++J, ++p)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 4 * (lane >> 4) + v, c = lane & 15;
                if (16 * I + r < n_f && 16 * J + c < n_f) {
                    double2 m = out[p * 256 + r * 16 + c];
                    m.x += double(RR[p][v]) + double(II[p][v]);
                    m.y += double(RI[p][v]) - double(IR[p][v]);
                    out[p * 256 + r * 16 + c] = m;
                }
            }
}

__global__ void munu_reset_kernel(unsigned int* cnt, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) cnt[i] = 0u;
}

template <typename F>
hipError_t dispatch_ns(int NS, F&& f) {
    switch (NS) {
        case 6: return f(std::integral_constant<int, 6>{});
        case 8: return f(std::integral_constant<int, 8>{});
        case 18: return f(std::integral_constant<int, 18>{});
        case 30: return f(std::integral_constant<int, 30>{});
        case 32: return f(std::integral_constant<int, 32>{});
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

int energy_blocks(int N) { return (N + kBlock - 1) / kBlock; }

__global__ void anneal_set_T_kernel(double* __restrict__ T_slot, int R, const double* __restrict__ sched,
                                    int n_sched, const unsigned long long* __restrict__ step) {
    const unsigned long long k = *step;
    if (k >= (unsigned long long)n_sched) return;
    const double T = sched[k];
    for (int s = threadIdx.x; s < R; s += blockDim.x) T_slot[s] = T;
}

__global__ void anneal_record_kernel(const double* __restrict__ e, int R, double* __restrict__ hist, int n_sched,
                                     unsigned long long* step, unsigned long long* pass_base,
                                     unsigned long long passes) {
    const unsigned long long k = *step;
    if (k < (unsigned long long)n_sched)
        for (int r = threadIdx.x; r < R; r += blockDim.x) hist[size_t(r) * n_sched + k] = e[r];
    __syncthreads();  // every thread has read *step
    if (threadIdx.x == 0) {
        *step = k + 1ull;
        *pass_base += passes;
    }
}

hipError_t launch_anneal_set_T(double* T_slot, int R, const double* sched, int n_sched,
                               const unsigned long long* step, hipStream_t s) {
    hipLaunchKernelGGL(anneal_set_T_kernel, dim3(1), dim3(kBlock), 0, s, T_slot, R, sched, n_sched, step);
    return hipGetLastError();
}

hipError_t launch_anneal_record(const double* e, int R, double* hist, int n_sched, unsigned long long* step,
                                unsigned long long* pass_base, unsigned long long passes, hipStream_t s) {
    hipLaunchKernelGGL(anneal_record_kernel, dim3(1), dim3(kBlock), 0, s, e, R, hist, n_sched, step, pass_base,
                       passes);
    return hipGetLastError();
}

hipError_t launch_energy_finish(const double* partial, int nblk, int n_cells, double* e,
                                const int32_t* slot_of, double* Esum, double* E2sum,
                                unsigned long long* nE, int R, hipStream_t s) {
    hipLaunchKernelGGL(energy_finish_kernel, dim3(unsigned(R)), dim3(kBlock), 0, s, partial, nblk,
                       double(n_cells), e, slot_of, Esum, E2sum, nE);
    return hipGetLastError();
}

hipError_t launch_reduce_accepts(const unsigned int* acc_wave, int n, unsigned long long* acc, int R,
                                 hipStream_t s) {
    if (R <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_accepts_kernel, dim3(unsigned(R)), dim3(kBlock), 0, s, acc_wave, n, acc);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_sweep_colour(const DeviceState<Real>& st, const ColourList& col, int kind,
                               unsigned long long pass_index, hipStream_t s) {
    const int spc = st.n_cells / (st.p0 * st.p1 * st.p2);
    const int bpc = (spc + kBlock - 1) / kBlock;
    const dim3 grid(unsigned(bpc * col.n_cls) * unsigned(st.R));  // replica-major, XCD-aware (xcd_split)
    return dispatch_ns(st.NS, [&](auto ns) -> hipError_t {
        constexpr int NS = decltype(ns)::value;
        auto go = [&](auto iso) -> hipError_t {
            constexpr bool ISO = decltype(iso)::value;
            switch (kind) {
                case 0: hipLaunchKernelGGL((sweep_colour_kernel<Real, NS, ISO, 0>), grid, dim3(kBlock), 0, s, st, col.cls, spc, bpc, pass_index); break;
                case 1: hipLaunchKernelGGL((sweep_colour_kernel<Real, NS, ISO, 1>), grid, dim3(kBlock), 0, s, st, col.cls, spc, bpc, pass_index); break;
                case 2: hipLaunchKernelGGL((sweep_colour_kernel<Real, NS, ISO, 2>), grid, dim3(kBlock), 0, s, st, col.cls, spc, bpc, pass_index); break;
                case 3: hipLaunchKernelGGL((sweep_colour_kernel<Real, NS, ISO, 3>), grid, dim3(kBlock), 0, s, st, col.cls, spc, bpc, pass_index); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        };
        return st.iso ? go(std::true_type{}) : go(std::false_type{});
    });
}

template <typename Real>
hipError_t launch_sweep_resident(const DeviceState<Real>& st, const int32_t* all_cls,
                                 const int32_t* colour_begin, int n_colours, int kind,
                                 int n_or, int n_sweeps, unsigned long long pass_index,
                                 hipStream_t s) {
    constexpr int BS = 512;
    const dim3 grid(unsigned(st.R));
    return dispatch_ns(st.NS, [&](auto ns) -> hipError_t {
        constexpr int NS = decltype(ns)::value;
        auto go = [&](auto iso) -> hipError_t {
            constexpr bool ISO = decltype(iso)::value;
            switch (kind) {
                case 0: hipLaunchKernelGGL((sweep_resident_kernel<Real, NS, ISO, 0, BS>), grid, dim3(BS), 0, s, st, all_cls, colour_begin, n_colours, n_or, n_sweeps, pass_index); break;
                case 1: hipLaunchKernelGGL((sweep_resident_kernel<Real, NS, ISO, 1, BS>), grid, dim3(BS), 0, s, st, all_cls, colour_begin, n_colours, n_or, n_sweeps, pass_index); break;
                case 2: hipLaunchKernelGGL((sweep_resident_kernel<Real, NS, ISO, 2, BS>), grid, dim3(BS), 0, s, st, all_cls, colour_begin, n_colours, n_or, n_sweeps, pass_index); break;
                case 3: hipLaunchKernelGGL((sweep_resident_kernel<Real, NS, ISO, 3, BS>), grid, dim3(BS), 0, s, st, all_cls, colour_begin, n_colours, n_or, n_sweeps, pass_index); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        };
        return st.iso ? go(std::true_type{}) : go(std::false_type{});
    });
}

template <typename Real>
hipError_t launch_fields(const DeviceState<Real>& st, int r, Real* out, hipStream_t s) {
    const dim3 grid(unsigned((st.N + kBlock - 1) / kBlock));
    return dispatch_ns(st.NS, [&](auto ns) -> hipError_t {
        constexpr int NS = decltype(ns)::value;
        if (st.iso) hipLaunchKernelGGL((fields_kernel<Real, NS, true>), grid, dim3(kBlock), 0, s, st, r, out);
        else hipLaunchKernelGGL((fields_kernel<Real, NS, false>), grid, dim3(kBlock), 0, s, st, r, out);
        return hipGetLastError();
    });
}

template <typename Real>
hipError_t launch_energy(const DeviceState<Real>& st, double* partial, double* e, double* Esum,
                         double* E2sum, unsigned long long* nE, hipStream_t s) {
    const int nblk = energy_blocks(st.N);
    const dim3 grid(unsigned(nblk) * unsigned(st.R));  // XCD-aware (xcd_split)
    hipError_t err = dispatch_ns(st.NS, [&](auto ns) -> hipError_t {
        constexpr int NS = decltype(ns)::value;
        if (st.iso) hipLaunchKernelGGL((energy_partial_kernel<Real, NS, true>), grid, dim3(kBlock), 0, s, st, partial);
        else hipLaunchKernelGGL((energy_partial_kernel<Real, NS, false>), grid, dim3(kBlock), 0, s, st, partial);
        return hipGetLastError();
    });
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(energy_finish_kernel, dim3(unsigned(st.R)), dim3(kBlock), 0, s, partial,
                       nblk, double(st.n_cells), e, st.slot_of, Esum, E2sum, nE);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_randomize(const DeviceState<Real>& st, uint32_t k0, uint32_t k1,
                            hipStream_t s) {
    const dim3 grid(unsigned((st.N + kBlock - 1) / kBlock), unsigned(st.R));
    hipLaunchKernelGGL(randomize_kernel<Real>, grid, dim3(kBlock), 0, s, st, k0, k1);
    return hipGetLastError();
}

hipError_t launch_exchange(const ExchangeArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(exchange_kernel, dim3(1), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bump(unsigned long long* ctr, unsigned long long n, hipStream_t s) {
    hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(1), 0, s, ctr, n);
    return hipGetLastError();
}

hipError_t launch_map_slots(const int32_t* slot_of, const int32_t* map, int32_t* out, int R,
                            hipStream_t s) {
    hipLaunchKernelGGL(map_slots_kernel, dim3(1), dim3(kBlock), 0, s, slot_of, map, out, R);
    return hipGetLastError();
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void fill_field_kernel(Real* __restrict__ B, int R, Real bx, Real by, Real bz) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r < R) {
        B[3 * r] = bx;
        B[3 * r + 1] = by;
        B[3 * r + 2] = bz;
    }
}

template <typename Real>
hipError_t launch_fill_field(Real* B, int R, double bx, double by, double bz, hipStream_t s) {
    hipLaunchKernelGGL(fill_field_kernel<Real>, dim3(unsigned((R + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, B, R,
                       Real(bx), Real(by), Real(bz));
    return hipGetLastError();
}
template hipError_t launch_fill_field<float>(float*, int, double, double, double, hipStream_t);
template hipError_t launch_fill_field<double>(double*, int, double, double, double, hipStream_t);

hipError_t launch_zero(void* p, size_t n_words32, hipStream_t s) {
    if (n_words32 == 0) return hipSuccess;
    size_t blocks = (n_words32 + kBlock - 1) / kBlock;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(zero_kernel, dim3(unsigned(blocks)), dim3(kBlock), 0, s,
                       static_cast<uint32_t*>(p), n_words32);
    return hipGetLastError();
}

// Whole 3-D DFT of one (replica, component, sublattice) field in LDS for
// cubic lattices of edge L = 4, 8, 12 or 16.  The field (L^3 reals) is read
// once, transformed along axis 2, 1, 0 in place and written once as the
// kept half F[r][a][mu][kh] (k2 < L / 2 + 1; the other K are the conjugates
// of their partners -K, since the field is real).  A thread owns one line
// per pass: it keeps the line's L values in registers, transforms them with
// dft_line (one mixed-radix step over 4-point DFTs, compile-time twiddles;
// the first pass on a real line), then writes them back over its own line,
// so lines never overlap.
// a thread per line: L^2 lines rounded up to whole waves
template <int L>
constexpr int dft3_threads() { return (L * L + 63) / 64 * 64; }

// Real fields out[r][a][i] = S[r][i][a] (N sites per replica) for the
// per-axis DFT path; one thread per output value.
template <typename Real>
__global__ __launch_bounds__(kBlock) void deinterleave_kernel(const Real* __restrict__ S, Real* __restrict__ out,
                                                              long long N, long long total) {
    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (e >= total) return;
    const long long ra = e / N, i = e - ra * N, r = ra / 3, a = ra - 3 * r;
    out[e] = S[(r * N + i) * 3 + a];
}

// Compile-time twiddles e^{-2 pi i m / L} for the whole-field DFT: Taylor
// series in double precision (exact 0 and +-1 at the quarter turns, so the
// compiler folds those products away).
namespace dftc {
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double taylor_sin(double x) {
    double t = x, s = x;
    for (int n = 1; n < 24; ++n) {
        t *= -x * x / double((2 * n) * (2 * n + 1));
        s += t;
    }
    return s;
}
constexpr double taylor_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int n = 1; n < 24; ++n) {
        t *= -x * x / double((2 * n - 1) * (2 * n));
        s += t;
    }
    return s;
}
// real and imaginary parts of e^{-2 pi i m / L}
constexpr double wre(int m, int L) {
    m = ((m % L) + L) % L;
    if ((4 * m) % L == 0) return (4 * m) / L == 0 ? 1.0 : ((4 * m) / L == 2 ? -1.0 : 0.0);
    const int mc = 2 * m > L ? m - L : m;  // angle in (-pi, pi]
    return taylor_cos(kTwoPi * double(mc) / double(L));
}
constexpr double wim(int m, int L) {
    m = ((m % L) + L) % L;
    if ((4 * m) % L == 0) return (4 * m) / L == 1 ? -1.0 : ((4 * m) / L == 3 ? 1.0 : 0.0);
    const int mc = 2 * m > L ? m - L : m;
    return -taylor_sin(kTwoPi * double(mc) / double(L));
}
}  // namespace dftc

// In-register DFT of P = 2, 3 or 4 points (e^{-2 pi i n k / P}).
template <typename Real, int P>
__device__ __forceinline__ void dft_small(Real (&r)[P], Real (&i)[P]) {
    if constexpr (P == 2) {
        const Real ar = r[0] + r[1], ai = i[0] + i[1];
        r[1] = r[0] - r[1];
        i[1] = i[0] - i[1];
        r[0] = ar;
        i[0] = ai;
    } else if constexpr (P == 3) {
        constexpr Real h = Real(0.86602540378443864676);  // sqrt(3) / 2
        const Real sr = r[1] + r[2], si = i[1] + i[2], dr = r[1] - r[2], di = i[1] - i[2];
        const Real mr = r[0] - Real(0.5) * sr, mi = i[0] - Real(0.5) * si;
        r[0] += sr;
        i[0] += si;
        r[1] = mr + h * di;
        i[1] = mi - h * dr;
        r[2] = mr - h * di;
        i[2] = mi + h * dr;
    } else {
        static_assert(P == 4, "2, 3 or 4 points");
        const Real ar = r[0] + r[2], ai = i[0] + i[2], br = r[0] - r[2], bi = i[0] - i[2];
        const Real cr = r[1] + r[3], ci = i[1] + i[3], er = r[1] - r[3], ei = i[1] - i[3];
        r[0] = ar + cr;
        i[0] = ai + ci;
        r[2] = ar - cr;
        i[2] = ai - ci;
        r[1] = br + ei;
        i[1] = bi - er;
        r[3] = br - ei;
        i[3] = bi + er;
    }
}

// In-register DFT of L = 4, 8, 12 or 16 points by one mixed-radix step,
// L = P Q with P = 4 (L = 4: P = 4, Q = 1): n = Q n1 + n2, k = k1 + P k2,
//   Z[n2][k1] = DFT_P over n1 of x[Q n1 + n2],  Z *= w_L^(n2 k1),
//   X[k1 + P k2] = DFT_Q over n2 of Z[n2][k1],
// with compile-time twiddles.
template <typename Real, int L>
__device__ __forceinline__ void dft_line(Real (&r)[L], Real (&i)[L]) {
    constexpr int P = 4, Q = L / 4;
    static_assert(L % 4 == 0 && Q >= 1 && Q <= 4, "4, 8, 12 or 16 points");
    Real zr[Q][P], zi[Q][P];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) {
            zr[n2][n1] = r[Q * n1 + n2];
            zi[n2][n1] = i[Q * n1 + n2];
        }
        dft_small<Real, P>(zr[n2], zi[n2]);
    }
    if constexpr (Q == 1) {
#pragma unroll
        for (int k = 0; k < L; ++k) {
            r[k] = zr[0][k];
            i[k] = zi[0][k];
        }
    } else {
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
            Real ur[Q], ui[Q];
#pragma unroll
            for (int n2 = 0; n2 < Q; ++n2) {
                const Real c = Real(dftc::wre(n2 * k1, L)), d = Real(dftc::wim(n2 * k1, L));
                ur[n2] = zr[n2][k1] * c - zi[n2][k1] * d;
                ui[n2] = zr[n2][k1] * d + zi[n2][k1] * c;
            }
            if constexpr (Q == 2 || Q == 3 || Q == 4) dft_small<Real, Q>(ur, ui);
#pragma unroll
            for (int k2 = 0; k2 < Q; ++k2) {
                r[k1 + P * k2] = ur[k2];
                i[k1 + P * k2] = ui[k2];
            }
        }
    }
}

template <typename Real, int L>
__global__ __launch_bounds__(dft3_threads<L>()) void dft3_kernel(const Real* __restrict__ spins,
                                                                 Real* __restrict__ F,
                                                                 const Real* __restrict__ tw, int n_sl) {
    // passes 1 and 2 and the output cover the kept half k2 < H only
    constexpr int NC = L * L * L, H = L / 2 + 1, NCH = L * L * H, BS = dft3_threads<L>();
    __shared__ Real x[2 * NC];
    (void)tw;  // twiddles are compile-time constants (dft_line)
    const size_t f_off = (size_t(blockIdx.y) * 3 * n_sl + size_t(blockIdx.x)) * size_t(NCH);
    // field (component a, sublattice mu) of spins S[r][i][a]: every third
    // value from 3 (mu NC) + a; the three components' workgroups are n_sl
    // apart in the grid, on one XCD when n_sl is a multiple of 8
    const int a = int(blockIdx.x) / n_sl, mu = int(blockIdx.x) - a * n_sl;
    const Real* src = spins + (size_t(blockIdx.y) * n_sl + mu) * 3 * size_t(NC) + a;
    {
        // every load of the thread issued before any value is stored, so
        // that they are in flight together
        constexpr int KN = (NC + BS - 1) / BS;
        Real v[KN];
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            const int i = int(threadIdx.x) + k * BS;
            v[k] = i < NC ? src[3 * i] : Real(0);
        }
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            const int i = int(threadIdx.x) + k * BS;
            if (i < NC) {
                x[2 * i] = v[k];
                x[2 * i + 1] = Real(0);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int pass = 0; pass < 3; ++pass) {
        // pass 0: rows (n0, n1) along n2; pass 1: (n0, k2 < H) along n1;
        // pass 2: (k1, k2 < H) along n0
        const int sa = pass == 0 ? 1 : (pass == 1 ? L : L * L);
        const int lines = pass == 0 ? L * L : L * H;
        for (int line = threadIdx.x; line < lines; line += BS) {
            const int base = pass == 0 ? line * L
                                       : (pass == 1 ? (line / H) * (L * L) + line % H : (line / H) * L + line % H);
            Real xr[L], xi[L];
#pragma unroll
            for (int j = 0; j < L; ++j) {
                xr[j] = x[2 * (base + j * sa)];
                // the first pass transforms real spins: a literal zero lets
                // the compiler drop the imaginary half of its arithmetic
                xi[j] = pass == 0 ? Real(0) : x[2 * (base + j * sa) + 1];
            }
            dft_line<Real, L>(xr, xi);
#pragma unroll
            for (int k = 0; k < L; ++k) {
                x[2 * (base + k * sa)] = xr[k];
                x[2 * (base + k * sa) + 1] = xi[k];
            }
        }
        __syncthreads();
    }
    Real* dst = F + 2 * f_off;
    for (int e = threadIdx.x; e < NCH; e += BS) {
        const int src = (e / H) * L + e % H;
        dst[2 * e] = x[2 * src];
        dst[2 * e + 1] = x[2 * src + 1];
    }
}

// One axis of the per-sublattice DFT: the LDS line kernel for La <= 64,
// the direct kernel otherwise.
template <typename Real, bool REAL_IN>
hipError_t dft_axis(const Real* in, Real* out, const Real* tw, int La, int stride, long long total,
                    hipStream_t s) {
    if (La > 64 || total >= (1ll << 30)) {
        const unsigned blocks = unsigned((total + kBlock - 1) / kBlock);
        hipLaunchKernelGGL((dft_axis_kernel<Real, REAL_IN>), dim3(blocks), dim3(kBlock), 0, s, in, out, tw, La,
                           stride, total);
        return hipGetLastError();
    }
    const unsigned n_cols = unsigned(total / La);
    const dim3 grid((n_cols + 63u) / 64u);
    if constexpr (std::is_same_v<Real, float>) {
        // FP32, 32, 48 or 64 points: one Cooley-Tukey step over 16-point GEMMs
        if (La == 32 || La == 48 || La == 64) {
            auto ct = [&](auto m) -> hipError_t {
                constexpr int M = decltype(m)::value;
                if (!REAL_IN && stride > 1)
                    hipLaunchKernelGGL((dft_ct_strided_kernel<M>), grid, dim3(256), 0, s, in, out, tw, stride, n_cols);
                else
                    hipLaunchKernelGGL((dft_ct_kernel<REAL_IN, M>), grid, dim3(256), 0, s, in, out, tw, stride, n_cols);
                return hipGetLastError();
            };
            if (La == 32) return ct(std::integral_constant<int, 2>{});
            if (La == 48) return ct(std::integral_constant<int, 3>{});
            return ct(std::integral_constant<int, 4>{});
        }
        // other lengths: the DFT as one complex GEMM on MFMA
        const int Kp = (La + 3) & ~3, MT = (La + 15) / 16;
        const size_t lds = (size_t(2) * Kp * 80 + size_t(2) * Kp * 16 * MT) * sizeof(float);
        auto go = [&](auto mt) -> hipError_t {
            constexpr int M = decltype(mt)::value;
            auto* fn = dft_mfma_kernel<REAL_IN, M>;
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(fn, grid, dim3(256), lds, s, in, out, tw, La, stride, n_cols);
            return hipGetLastError();
        };
        switch (MT) {
            case 1: return go(std::integral_constant<int, 1>{});
            case 2: return go(std::integral_constant<int, 2>{});
            case 3: return go(std::integral_constant<int, 3>{});
            default: return go(std::integral_constant<int, 4>{});
        }
    }
    const size_t lds = (size_t(2) * 64 * La + 2 * La) * sizeof(Real);
    auto go = [&](auto kpt) -> hipError_t {
        constexpr int KPT = decltype(kpt)::value;
        auto* fn = dft_lds_kernel<Real, REAL_IN, KPT>;
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(fn, grid, dim3(256), lds, s, in, out, tw, La, stride, n_cols);
        return hipGetLastError();
    };
    if (La <= 16) return go(std::integral_constant<int, 4>{});
    if (La <= 32) return go(std::integral_constant<int, 8>{});
    return go(std::integral_constant<int, 16>{});
}

static bool dft3_applies(int L0, int L1, int L2) {
    return L0 == L1 && L1 == L2 && (L0 == 4 || L0 == 8 || L0 == 12 || L0 == 16);
}
static bool ct_length(int L) { return L == 32 || L == 48 || L == 64; }

int dft_kept_k2(int real_bytes, int L0, int L1, int L2, long long fields_per_replica_total) {
    if (dft3_applies(L0, L1, L2)) return L2 / 2 + 1;
    if (real_bytes == 4 && ct_length(L1) && ct_length(L2) && fields_per_replica_total < (1ll << 30))
        return L2 / 2 + 1;
    return L2;
}

template <typename Real>
hipError_t launch_dft(const SsfState<Real>& st, int r_begin, int r_count, hipStream_t s) {
    const size_t nc = size_t(st.n_cells);
    const long long total = (long long)r_count * 3 * st.n_sl * (long long)nc;
    const Real* spins = st.S + size_t(r_begin) * 3 * st.n_sl * nc;
    if (st.H != dft_kept_k2(int(sizeof(Real)), st.L0, st.L1, st.L2, 3ll * st.n_sl * (long long)nc) ||
        st.n_kh != st.L0 * st.L1 * st.H)
        return hipErrorInvalidValue;
    if (dft3_applies(st.L0, st.L1, st.L2)) {
        const dim3 grid(unsigned(3 * st.n_sl), unsigned(r_count));
        auto go = [&](auto l) -> hipError_t {
            constexpr int LL = decltype(l)::value;
            hipLaunchKernelGGL((dft3_kernel<Real, LL>), grid, dim3(dft3_threads<LL>()), 0, s, spins, st.F,
                               st.tw0, st.n_sl);
            return hipGetLastError();
        };
        switch (st.L0) {
            case 4: return go(std::integral_constant<int, 4>{});
            case 8: return go(std::integral_constant<int, 8>{});
            case 12: return go(std::integral_constant<int, 12>{});
            case 16: return go(std::integral_constant<int, 16>{});
            default: break;
        }
    }
    if constexpr (std::is_same_v<Real, float>) {
        // axes 2 and 1 of every plane in one pass (spins -> G, the kept half
        // k2 < H), then axis 0 over those columns: G -> F
        auto plane = [&](auto m1, auto m2) -> hipError_t {
            constexpr int M1 = decltype(m1)::value, M2 = decltype(m2)::value;
            const long long planes = total / ((long long)st.L1 * st.L2);
            hipLaunchKernelGGL((dft_plane_kernel<M1, M2>), dim3(unsigned(planes)), dim3(256), 0, s, spins, st.G,
                               st.tw1, st.tw2, st.n_sl, st.L0);
            hipError_t e = hipGetLastError();
            const long long total_h = planes * st.L1 * st.H;
            if (e == hipSuccess) e = dft_axis<Real, false>(st.G, st.F, st.tw0, st.L0, st.L1 * st.H, total_h, s);
            return e;
        };
        auto m2 = [&](auto m1) -> hipError_t {
            switch (st.L2) {
                case 32: return plane(m1, std::integral_constant<int, 2>{});
                case 48: return plane(m1, std::integral_constant<int, 3>{});
                default: return plane(m1, std::integral_constant<int, 4>{});
            }
        };
        if (st.H < st.L2) {  // the plane path (dft_kept_k2)
            switch (st.L1) {
                case 32: return m2(std::integral_constant<int, 2>{});
                case 48: return m2(std::integral_constant<int, 3>{});
                default: return m2(std::integral_constant<int, 4>{});
            }
        }
    }
    if (st.H != st.L2) return hipErrorInvalidValue;
    // spins S[r][i][a] -> real fields G[r][a][mu][cell]; axis 2 (stride 1):
    // G -> F; axis 1 (stride L2): F -> G; axis 0: G -> F
    hipLaunchKernelGGL((deinterleave_kernel<Real>), dim3(unsigned((total + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       spins, st.G, (long long)st.n_sl * (long long)nc, total);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = dft_axis<Real, true>(st.G, st.F, st.tw2, st.L2, 1, total, s);
    if (e == hipSuccess) e = dft_axis<Real, false>(st.F, st.G, st.tw1, st.L1, st.L2, total, s);
    if (e == hipSuccess) e = dft_axis<Real, false>(st.G, st.F, st.tw0, st.L0, st.L1 * st.L2, total, s);
    return e;
}

template <typename Real>
hipError_t launch_contract(const SsfState<Real>& st, const CorrArgs& c, int r_begin, int r_count,
                           hipStream_t s) {
    if constexpr (std::is_same_v<Real, float>) {
        // FP32: fields through LDS, the fold split across waves
        const int T = (3 * st.n_sl + 15) / 16;
        const dim3 g64(unsigned((st.n_kh + 63) / 64), unsigned(r_count));
        switch (T) {
            case 1: hipLaunchKernelGGL((contract_stage_kernel<1, false>), g64, dim3(kBlock), 0, s, st, c, r_begin, nullptr, nullptr, 0, nullptr); return hipGetLastError();
            case 2: hipLaunchKernelGGL((contract_stage_kernel<2, false>), g64, dim3(kBlock), 0, s, st, c, r_begin, nullptr, nullptr, 0, nullptr); return hipGetLastError();
            case 3: hipLaunchKernelGGL((contract_stage_kernel<3, false>), g64, dim3(kBlock), 0, s, st, c, r_begin, nullptr, nullptr, 0, nullptr); return hipGetLastError();
            default: break;
        }
    }
    const dim3 grid(unsigned((st.n_kh + kContractBlock - 1) / kContractBlock), unsigned(r_count));
    hipLaunchKernelGGL(contract_kernel<Real>, grid, dim3(kContractBlock), 0, s, st, c, r_begin);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_munu_stage(const SsfState<Real>& st, const int32_t* slot_of, int r_begin,
                             int r_count, const MunuArgs& m, hipStream_t s) {
    const dim3 grid(unsigned((st.n_kh + 63) / 64), unsigned(r_count));
    switch (m.T) {
        case 1: hipLaunchKernelGGL((munu_stage_kernel<Real, 1>), grid, dim3(kBlock), 0, s, st, slot_of, r_begin, m.buf, m.cnt, m.Bmax); break;
        case 2: hipLaunchKernelGGL((munu_stage_kernel<Real, 2>), grid, dim3(kBlock), 0, s, st, slot_of, r_begin, m.buf, m.cnt, m.Bmax); break;
        case 3: hipLaunchKernelGGL((munu_stage_kernel<Real, 3>), grid, dim3(kBlock), 0, s, st, slot_of, r_begin, m.buf, m.cnt, m.Bmax); break;
        default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(munu_count_kernel, dim3(1), dim3(kBlock), 0, s, slot_of, r_begin, r_count,
                       m.cnt, m.Bmax);
    return hipGetLastError();
}

hipError_t launch_contract_stage(const SsfState<float>& st, const CorrArgs& c, int r_begin, int r_count,
                                 const MunuArgs& m, hipStream_t s) {
    const dim3 grid(unsigned((st.n_kh + 63) / 64), unsigned(r_count));
    switch (m.T) {
        case 1: hipLaunchKernelGGL((contract_stage_kernel<1, true>), grid, dim3(kBlock), 0, s, st, c, r_begin, m.buf, m.cnt, m.Bmax, m.cbuf); break;
        case 2: hipLaunchKernelGGL((contract_stage_kernel<2, true>), grid, dim3(kBlock), 0, s, st, c, r_begin, m.buf, m.cnt, m.Bmax, m.cbuf); break;
        case 3: hipLaunchKernelGGL((contract_stage_kernel<3, true>), grid, dim3(kBlock), 0, s, st, c, r_begin, m.buf, m.cnt, m.Bmax, m.cbuf); break;
        default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(munu_count_kernel, dim3(1), dim3(kBlock), 0, s, c.slot_of, r_begin, r_count, m.cnt, m.Bmax);
    return hipGetLastError();
}

hipError_t launch_corr_flush(const MunuArgs& m, hipStream_t s) {
    if (!m.cbuf) return hipSuccess;
    const long long n = (long long)m.n_corr * m.n_k;
    hipLaunchKernelGGL(corr_flush_kernel, dim3(unsigned((n + kBlock - 1) / kBlock), unsigned(m.n_slots)),
                       dim3(kBlock), 0, s, m.cbuf, m.cnt, m.acc_s, m.acc_s2, m.n_corr, m.n_k, m.H, m.L2, m.n_k_full,
                       m.Bmax);
    return hipGetLastError();
}

hipError_t launch_munu_flush(const MunuArgs& m, hipStream_t s) {
    const dim3 grid(unsigned((m.n_k + 3) / 4), unsigned(m.n_slots));
    switch (m.T) {
        case 1: hipLaunchKernelGGL((munu_flush_kernel<1>), grid, dim3(256), 0, s, m.buf, m.cnt, m.acc, m.n_k, m.Bmax, m.n_f); break;
        case 2: hipLaunchKernelGGL((munu_flush_kernel<2>), grid, dim3(256), 0, s, m.buf, m.cnt, m.acc, m.n_k, m.Bmax, m.n_f); break;
        case 3: hipLaunchKernelGGL((munu_flush_kernel<3>), grid, dim3(256), 0, s, m.buf, m.cnt, m.acc, m.n_k, m.Bmax, m.n_f); break;
        default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(munu_reset_kernel, dim3(1), dim3(kBlock), 0, s, m.cnt, m.n_slots);
    return hipGetLastError();
}

}  // namespace pyro

// Explicit instantiations.
namespace pyro {
template hipError_t launch_sweep_colour<float>(const DeviceState<float>&, const ColourList&, int, unsigned long long, hipStream_t);
template hipError_t launch_sweep_colour<double>(const DeviceState<double>&, const ColourList&, int, unsigned long long, hipStream_t);
template hipError_t launch_sweep_resident<float>(const DeviceState<float>&, const int32_t*, const int32_t*, int, int, int, int, unsigned long long, hipStream_t);
template hipError_t launch_sweep_resident<double>(const DeviceState<double>&, const int32_t*, const int32_t*, int, int, int, int, unsigned long long, hipStream_t);
template hipError_t launch_fields<float>(const DeviceState<float>&, int, float*, hipStream_t);
template hipError_t launch_fields<double>(const DeviceState<double>&, int, double*, hipStream_t);
template hipError_t launch_energy<float>(const DeviceState<float>&, double*, double*, double*, double*, unsigned long long*, hipStream_t);
template hipError_t launch_energy<double>(const DeviceState<double>&, double*, double*, double*, double*, unsigned long long*, hipStream_t);
template hipError_t launch_randomize<float>(const DeviceState<float>&, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_randomize<double>(const DeviceState<double>&, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_dft<float>(const SsfState<float>&, int, int, hipStream_t);
template hipError_t launch_dft<double>(const SsfState<double>&, int, int, hipStream_t);
template hipError_t launch_contract<float>(const SsfState<float>&, const CorrArgs&, int, int, hipStream_t);
template hipError_t launch_contract<double>(const SsfState<double>&, const CorrArgs&, int, int, hipStream_t);
template hipError_t launch_munu_stage<float>(const SsfState<float>&, const int32_t*, int, int, const MunuArgs&, hipStream_t);
template hipError_t launch_munu_stage<double>(const SsfState<double>&, const int32_t*, int, int, const MunuArgs&, hipStream_t);
}  // namespace pyro
