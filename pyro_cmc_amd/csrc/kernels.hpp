// Launch interface between the context code (capi.cpp) and the HIP kernels
// (kernels.hip).  Everything here is host-callable; the kernels themselves
// stay private to kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "topology.hpp"

namespace pyro {

// Device view of a context's sweep state, shared by the sweep, field and
// energy kernels.  Spins are S[r][i][c], c = x, y, z: one site's three
// components are adjacent, so a neighbour gather is one 12-byte (FP32) load.
template <typename Real>
struct DeviceState {
    Real* S = nullptr;                   // [R][N][3]
    int8_t* dir = nullptr;               // [R][N]
    const int32_t* nbr = nullptr;        // [NS][N] slot-major neighbour table
    const Real* cpl = nullptr;           // isotropic [n_sl][NS] else [n_sl][NS][9]
    const double* T_slot = nullptr;      // [R] temperature of each slot
    const int32_t* slot_of = nullptr;    // [R] slot of each replica
    const Real* B = nullptr;             // [R][3] global field per replica
    // Accepted-update counting, only when a caller asks for the counts
    // (both NULL otherwise).  Whole-replica kernels add once per workgroup
    // to acc[r]; per-colour kernels keep one counter per wave slot,
    // acc_wave[r * acc_waves + slot], updated by its owning wave alone, so
    // no two waves contend on an address.  launch_reduce_accepts folds the
    // slots into acc.
    unsigned long long* acc = nullptr;   // [R]
    unsigned int* acc_wave = nullptr;    // [R][acc_waves]
    int acc_waves = 0;
    const unsigned long long* pass_base = nullptr;  // device pass counter (graph mode)
    uint32_t k0 = 0, k1 = 0;             // RNG key
    double T_ref = 1.0;
    int N = 0, n_cells = 0, R = 0, NS = 0;
    int L0 = 0, L1 = 0, L2 = 0;
    int p0 = 1, p1 = 1, p2 = 1;          // colour residue periods
    bool iso = true;
};

// One colour = a list of classes (sublattice, residues), packed as
// sl | r0 << 16 | r1 << 20 | r2 << 24.
struct ColourList {
    const int32_t* cls = nullptr;  // device array of class descriptors
    int n_cls = 0;
};

// Uniforms one update of each kind consumes from its block: Metropolis 5
// (two Box-Muller pairs and the acceptance), lifted 2 (the exponential
// displacement and the acceptance), lifted rotation 1, overrelaxation 1.
constexpr int kUniformsOf[4] = {5, 2, 1, 1};

// One pass of `kind` over one colour for every replica (one launch).
// pass_index is added to *pass_base when pass_base is set.
template <typename Real>
hipError_t launch_sweep_colour(const DeviceState<Real>& st, const ColourList& col, int kind,
                               unsigned long long pass_index, hipStream_t s);

// A whole sequence of passes (first kind, then n_or overrelaxation passes,
// repeated n_sweeps times) over all colours, one resident workgroup per
// replica synchronising with workgroup barriers between colours.
template <typename Real>
hipError_t launch_sweep_resident(const DeviceState<Real>& st, const int32_t* all_cls,
                                 const int32_t* colour_begin, int n_colours, int kind,
                                 int n_or, int n_sweeps, unsigned long long pass_index,
                                 hipStream_t s);

// Couplings of the pyrochlore-specialised kernels: shells = 1 (J1),
// 2 (J1, J2) or 3 (J1, J2, J3); Jx = J1 (Delta - 1) / 3 when xxz.
// per_replica, when set, holds (J1, J2, J3, Jx) for every replica and
// overrides the scalars (a batch of independent models in one context).
template <typename Real>
struct PyroCouplings {
    Real J1 = 0, J2 = 0, J3 = 0, Jx = 0;
    int shells = 1;
    bool xxz = false;
    const Real* per_replica = nullptr;  // [R][4] or null
    __device__ __forceinline__ PyroCouplings at(int r) const {
        PyroCouplings c = *this;
        if (per_replica) {
            c.J1 = per_replica[4 * r];
            c.J2 = per_replica[4 * r + 1];
            c.J3 = per_replica[4 * r + 2];
            c.Jx = per_replica[4 * r + 3];
        }
        return c;
    }
};
// One pass of `kind` over one colour (4 colours = pyro_sl when shells < 3,
// else 16 colours = sublattice) for every replica.  With epart (the last
// pass of a sweep, every colour in order): also each replica's energy as
// pyro_colour_energy_blocks(n_cells) partials epart[r][...], which
// launch_energy_finish sums once the pass's last colour has run.
// fused (the fused pass order, kind 0, 1 or 2): each site's first OR update
// follows its `kind` update in the same launch, from the same field, with
// the sweep's next uniform; ucomp (kind 3 only): the component of the
// block that an OR pass of the fused order running on its own draws.
template <typename Real>
hipError_t launch_pyro_colour(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int colour,
                              int kind, unsigned long long pass_index, hipStream_t s, double* epart = nullptr,
                              bool fused = false, int ucomp = 0);
__host__ __device__ constexpr int pyro_colour_energy_blocks(int n_cells) { return 16 * ((n_cells + 63) / 64); }
// Whole sweeps with one resident workgroup (1024 threads, two cells per
// thread) per replica; lattices of at most 2048 cells.  With epart the last
// pass also leaves the energy partials launch_pyro_colour leaves, bit for bit.
constexpr int kPyroResidentMaxCells = 2048;
// Whole sweeps with the per-colour work of launch_pyro_colour held on one
// XCD per replica in one launch, the colour stages separated by
// per-replica counters in that XCD's L2 (bar[32 (R + 8)]: R counters and 8
// XCD tickets, 128 bytes apart, zeroed by the launch), spins read past the
// L1: bit-identical to
// launch_pyro_colour pass after pass (epart and fused as for
// launch_pyro_resident).  Every wait is bounded by timeout_ticks of the
// device's wall clock from the workgroup's arrival at the previous stage; a
// timeout sets *err (then the results are invalid).  hipErrorNotSupported when the grid
// cannot be resident at once.
// Workgroups of one XCD-resident launch: 8 XCDs x ceil(R / 8) replicas x
// blocks per replica and colour stage.
inline int pyro_xcd_workgroups(int R, int n_cells, int shells) {
    const int bpr = ((n_cells + 63) / 64) * (shells == 3 ? 1 : 4);
    return 8 * ((R + 7) / 8) * bpr;
}
template <typename Real>
hipError_t launch_pyro_xcd(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int kind, int n_or,
                           int n_sweeps, unsigned long long pass_index, hipStream_t s, double* epart, bool fused,
                           unsigned int* bar, int* err, unsigned long long timeout_ticks);

// One context of a batch launch (launch_pyro_batch), in device memory.
template <typename Real>
struct BatchMember {
    DeviceState<Real> st;
    PyroCouplings<Real> J;
    double* epart = nullptr;             // energy partials as launch_pyro_colour leaves them (EN launches)
    int* err = nullptr;                  // the context's wait-timeout flag
    unsigned long long timeout = 0;      // the context's wait bound, in wall-clock ticks
    unsigned long long pass_index = 0;   // added to *st.pass_base
    int item0 = 0;                       // counter index of its replica 0 in the launch
};
// Whole sweeps of several contexts in one persistent launch: the members'
// replicas placed whole on XCDs by a work table of per_xcd int4 entries per
// XCD (member, replica, worker, workers; member < 0 idle), grid 8 per_xcd,
// one-wave workgroups; bar[32 (n_items + 8)] holds the n_items replica
// counters and 8 XCD tickets and is zeroed by the launch.  Every member
// has replicas below 2^24 sites and shares the template arguments (precision, coupling shells, per-replica
// couplings, kind, fused order, energy partials).  Each member's results
// equal launch_pyro_colour's pass after pass.  hipErrorNotSupported when
// the grid cannot be resident at once.
template <typename Real>
hipError_t launch_pyro_batch(const BatchMember<Real>* members, const int4* work, int per_xcd, int n_items, int shells,
                             bool per_rep, int kind, int n_or, int n_sweeps, bool energy, bool fused,
                             unsigned int* bar, hipStream_t s);
// Workgroups of launch_pyro_batch that can be resident at once on the
// current device (0 on error).
template <typename Real>
int pyro_batch_capacity(int shells, bool per_rep, int kind, bool energy, bool fused);
// fused: the fused pass order of launch_pyro_colour, bit for bit.
template <typename Real>
hipError_t launch_pyro_resident(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int kind,
                                int n_or, int n_sweeps, unsigned long long pass_index,
                                hipStream_t s, double* epart = nullptr, bool fused = false);

// Tiled sweeps: one call launches one of the 8 checkerboard phases of a
// sweep; its workgroups hold a tile + one-sided halo in LDS and run every
// pass.  Tile kinds (cells per axis): 1 = 4x4x4, 2 = 6x6x4, 3 = 8x4x4,
// 4 = 4x4x8, 5 = 8x8x4.  A kind applies when every L_a is divisible by twice
// its edge and its LDS fits.
constexpr int kPyroTileKinds = 5;
inline void pyro_tile_dims(int tile_kind, int dims[3]) {
    static const int T[kPyroTileKinds + 1][3] = {{0, 0, 0}, {4, 4, 4}, {6, 6, 4}, {8, 4, 4}, {4, 4, 8}, {8, 8, 4}};
    const int k = tile_kind >= 1 && tile_kind <= kPyroTileKinds ? tile_kind : 0;
    for (int a = 0; a < 3; ++a) dims[a] = T[k][a];
}
// Threads per tiled workgroup: the tile's cells rounded up to whole waves.
constexpr int pyro_tile_threads(int cells) { return cells >= 256 ? 256 : (cells + 63) / 64 * 64; }
// LDS bytes of a tiled workgroup with sublattice stride ls: spins of the
// tile + one-sided halo (3 x 16 x ls reals) and, for the lifted kinds, the
// interior's directions as bits (16 x ceil(T0 T1 T2 / 32) words).
constexpr size_t pyro_tile_lds_at(int ls, int T0, int T1, int T2, int real_size, bool lifted) {
    return size_t(3) * 16 * size_t(ls) * size_t(real_size) +
           (lifted ? size_t(16) * 4 * ((T0 * T1 * T2 + 31) / 32) : 0);
}
// Sublattice stride of a tiled workgroup's LDS arrays, in entries: the box
// of (T0+1)(T1+1)(T2+1) cells rounded up to a multiple of 64, so that one
// cell of every sublattice and component lies a multiple of 256 B from the
// others and the gathers pair up as ds_read2st64 from a few base addresses,
// unless the padding would cost a workgroup per CU (160 KB of LDS).
constexpr int pyro_tile_stride(int T0, int T1, int T2, int real_size, bool lifted) {
    const int bc = (T0 + 1) * (T1 + 1) * (T2 + 1), padded = (bc + 63) / 64 * 64;
    const size_t cu = 163840;
    return cu / pyro_tile_lds_at(padded, T0, T1, T2, real_size, lifted) ==
                   cu / pyro_tile_lds_at(bc, T0, T1, T2, real_size, lifted)
               ? padded
               : bc;
}
constexpr size_t pyro_tile_lds(int T0, int T1, int T2, int real_size, bool lifted) {
    return pyro_tile_lds_at(pyro_tile_stride(T0, T1, T2, real_size, lifted), T0, T1, T2, real_size, lifted);
}
// fused (pyro_set_fused_passes): within each colour stage a site takes its
// `kind` update and its first OR update back to back, from one field
// evaluation; further OR passes run pass by pass after them.  Else every
// pass runs pass by pass.
template <typename Real>
hipError_t launch_pyro_tiled(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int tile_kind,
                             int phase, int kind, int n_or, unsigned long long pass_index, hipStream_t s,
                             bool fused = false);

// Local fields (no global field) of replica r into out[3][N].
template <typename Real>
hipError_t launch_fields(const DeviceState<Real>& st, int r, Real* out, hipStream_t s);

// Energy per cell of every replica -> e[R]; partial[R][nblk] is scratch.
// If slot sums are given, also adds e and e^2 to the replica's slot.
template <typename Real>
hipError_t launch_energy(const DeviceState<Real>& st, double* partial, double* e,
                         double* Esum, double* E2sum, unsigned long long* nE, hipStream_t s);
int energy_blocks(int N);
// Partials per replica of launch_pyro_energy (64-site blocks).
inline int pyro_energy_blocks(int N) { return (N + 63) / 64; }
// Stencil form of the partial sums for contexts with pyrochlore J1-J2-J3
// couplings: the same partial[R][energy_blocks(N)] layout, no table read.
template <typename Real>
hipError_t launch_pyro_energy(const DeviceState<Real>& st, const PyroCouplings<Real>& J, double* partial,
                              hipStream_t s);
// Tiled form of launch_pyro_energy for a tile kind of pyro_tile_dims whose
// edges divide the lattice: one workgroup per (tile, replica) reads the tile
// and its one-sided halo into LDS once; partial[R][*nblk], *nblk = tiles per
// replica.  Each site's term is computed as launch_pyro_energy computes it.
template <typename Real>
hipError_t launch_pyro_energy_tiled(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int tile_kind,
                                    double* partial, int* nblk, hipStream_t s);
// Local field (without the global field) of one replica, out[3][N], with
// the stencil field.
template <typename Real>
hipError_t launch_pyro_fields(const DeviceState<Real>& st, const PyroCouplings<Real>& J, int replica, Real* out,
                              hipStream_t s);
// Anneal bookkeeping on the device, replayable from a graph:
// T_slot[s] = sched[*step] for s < R ...
hipError_t launch_anneal_set_T(double* T_slot, int R, const double* sched, int n_sched,
                               const unsigned long long* step, hipStream_t s);
// ... and hist[r][*step] = e[r], *step += 1, *pass_base += passes.
hipError_t launch_anneal_record(const double* e, int R, double* hist, int n_sched, unsigned long long* step,
                                unsigned long long* pass_base, unsigned long long passes, hipStream_t s);
// Second stage of launch_energy (sums the partials of every replica).
hipError_t launch_energy_finish(const double* partial, int nblk, int n_cells, double* e,
                                const int32_t* slot_of, double* Esum, double* E2sum,
                                unsigned long long* nE, int R, hipStream_t s);

// Uniform-on-sphere spins, lifted directions +1.
template <typename Real>
hipError_t launch_randomize(const DeviceState<Real>& st, uint32_t k0, uint32_t k1, hipStream_t s);

// Replica exchange of slot pairs (s, s+1), s = round parity mod 2; uses the
// per-replica energies e[R] (per cell).  Increments *round and adds
// passes_per_round to *pass_base when pass_base is non-null.
struct ExchangeArgs {
    const double* e = nullptr;
    int32_t* rep_of = nullptr;
    int32_t* slot_of = nullptr;
    const double* T_slot = nullptr;
    unsigned long long* att = nullptr;
    unsigned long long* accp = nullptr;
    unsigned long long* round = nullptr;
    unsigned long long* pass_base = nullptr;
    unsigned long long passes_per_round = 0;
    unsigned long long every = 1;  // exchange after every `every`-th round
    double n_cells = 1.0;
    uint32_t k0 = 0, k1 = 0;
    int R = 0;
};
hipError_t launch_exchange(const ExchangeArgs& a, hipStream_t s);

// Adds n to *ctr on the device (graph-safe counter bump).
hipError_t launch_bump(unsigned long long* ctr, unsigned long long n, hipStream_t s);
// out[r] = map[slot_of[r]] for r < R.
hipError_t launch_map_slots(const int32_t* slot_of, const int32_t* map, int32_t* out, int R,
                            hipStream_t s);
// Sets n words to zero (graph-safe replacement for a memset).
hipError_t launch_zero(void* p, size_t n_words32, hipStream_t s);
// B[r][0..2] = (bx, by, bz) for r < R (a field from kernel arguments, graph-safe).
template <typename Real>
hipError_t launch_fill_field(Real* B, int R, double bx, double by, double bz, hipStream_t s);
// acc[r] += sum of the n per-wave counters acc_wave[r][0..n), r < R.
hipError_t launch_reduce_accepts(const unsigned int* acc_wave, int n, unsigned long long* acc, int R,
                                 hipStream_t s);

// ---------------------------------------------------------------------------
// Static structure factor.
// ---------------------------------------------------------------------------
template <typename Real>
// The fields are real, so A(-K) = conj(A(K)): the fast DFT paths keep only
// k2 < H = L2 / 2 + 1 (the kept half, n_kh = L0 L1 H values per field, kept
// index kh = (k0 L1 + k1) H + k2); H = L2 keeps every K.  Accumulators of the
// contracted correlators are indexed by the full K: the contraction writes
// the kept K and, where the centred q of -K is not minus that of K (a
// Nyquist index on axis 0 or 1), -K too; the rest is conj(S(-K)) on read.
struct SsfState {
    const Real* S = nullptr;         // [R][N][3]
    Real* F = nullptr;               // [R][3][n_sl][n_kh] complex (2 Real)
    Real* G = nullptr;               // same, scratch
    const Real* tw0 = nullptr;       // twiddles per axis, [L_a] complex
    const Real* tw1 = nullptr;
    const Real* tw2 = nullptr;
    const Real* phase = nullptr;     // [n_sl][n_cells] complex e^{-i q(K).r_mu}, full K
    int R = 0, n_sl = 0, n_cells = 0, L0 = 0, L1 = 0, L2 = 0;
    int H = 0, n_kh = 0;             // kept k2 per (k0, k1) line, kept values per field
};

// k2 values per (k0, k1) line that launch_dft produces for a lattice (L2 / 2
// + 1 on the whole-field and plane paths, else L2); fields_per_replica_total
// = 3 n_sl n_cells.
int dft_kept_k2(int real_bytes, int L0, int L1, int L2, long long fields_per_replica_total);

// Per-sublattice DFT of all replicas' spins: F[r][a][mu][kh] (complex), the
// kept K of st.H.
template <typename Real>
hipError_t launch_dft(const SsfState<Real>& st, int r_begin, int r_count, hipStream_t s);

// Contracted correlators of the transformed replicas, accumulated into
// acc[slot][c][K][2] (and acc2 if non-null); n_samples[slot] += 1.
struct CorrArgs {
    const int32_t* slot_of = nullptr;   // [R] slot per replica (<0: skip)
    const int32_t* alpha = nullptr;     // [n_corr]
    const int32_t* beta = nullptr;      // [n_corr]
    double* acc = nullptr;              // [n_slots][n_corr][n_k][2]
    double* acc2 = nullptr;             // same or null
    unsigned long long* n_samples = nullptr;  // [n_slots]
    int n_corr = 0;
};
template <typename Real>
hipError_t launch_contract(const SsfState<Real>& st, const CorrArgs& c, int r_begin,
                           int r_count, hipStream_t s);

// Sublattice-resolved correlators.  The transformed fields of each sampled
// replica are staged as one FP32 row X = [Re a | Im a] (a_f, f = a*n_sl+mu,
// each half padded to 16*T) into buf[slot][k][cnt[slot]][32*T]; a flush
// accumulates X^T X per (slot, k) on FP32 MFMA and adds
// Re M = RR + II, Im M = RI - IR into the FP64 accumulators, M = sum
// conj(a)^T a, then clears the counts.  M is Hermitian, so only its 16 x 16
// tile pairs (I, J) with I <= J are kept: acc[slot][k][p][16][16][2] with
// p = munu_acc_pair(I, J, T) (element (f, g) of a lower tile is the
// conjugate of (g, f)).
inline int munu_acc_pairs(int T) { return T * (T + 1) / 2; }
inline int munu_acc_pair(int I, int J, int T) { return I * T - I * (I - 1) / 2 + (J - I); }
struct MunuArgs {
    float* buf = nullptr;
    unsigned int* cnt = nullptr;   // [n_slots]
    double* acc = nullptr;
    int Bmax = 0, n_k = 0, n_slots = 0, n_f = 0, T = 0;  // n_k: the kept K (SsfState::n_kh)
    // buffered contracted correlators of the FP32 staging path (null: the
    // contraction updates its FP64 sums directly): cbuf[slot][b][c][kh],
    // added to acc_s / acc_s2 ([slot][c][n_k_full][2]) by the flush
    float2* cbuf = nullptr;
    double* acc_s = nullptr;
    double* acc_s2 = nullptr;
    int n_corr = 0, H = 0, L2 = 0, n_k_full = 0;
};
template <typename Real>
hipError_t launch_munu_stage(const SsfState<Real>& st, const int32_t* slot_of, int r_begin,
                             int r_count, const MunuArgs& m, hipStream_t s);
// launch_contract and launch_munu_stage of FP32 contexts in one kernel (one
// read of the transformed fields), with the same results.
hipError_t launch_contract_stage(const SsfState<float>& st, const CorrArgs& c, int r_begin, int r_count,
                                 const MunuArgs& m, hipStream_t s);
hipError_t launch_munu_flush(const MunuArgs& m, hipStream_t s);
// The buffered contracted samples (MunuArgs::cbuf) into the FP64 sums, in
// sample order; before launch_munu_flush, which resets the row counts.
hipError_t launch_corr_flush(const MunuArgs& m, hipStream_t s);


// Luttinger-Tisza spectrum over the L[0] x L[1] x L[2] grid of k with
// q_a = 2 pi K'_a / (cell_a L_a), K' the centred index: builds the
// (3 n_sl)^2 Hermitian M(k) = sum_bonds J e^{i q.d} on the device and
// finds its lowest eigenvalue there (Householder tridiagonalisation and
// Sturm multisection per k); the eigenvector at k* is solved on the host.
// n_sl must be 16.  Writes lambda_min[k] (flat order
// (K0 L1 + K1) L2 + K2), the first k of the strict minimum into k_star and
// the lowest eigenvector there into eigvec[3 n_sl][2].  Returns an empty
// string or the error.
std::string lt_spectrum(const LtBondTable& table, int n_sl, const int L[3], const int cell[3],
                        double* lambda_min, int32_t* k_star, double* eigvec);

}  // namespace pyro
