// C ABI of libpyro_cmc.so: context management, host<->device layout
// conversion, launch sequencing (eager or HIP-graph), autotuning of the sweep
// launch strategy and kernel timing.  Declarations and the reference
// interface each entry point replaces are in include/pyro_cmc.h.

#include "pyro_cmc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "rng.hpp"
#include "topology.hpp"

namespace {

thread_local std::string g_last_error;

// Tile kinds of the tiled strategy in order of preference: the first that
// applies to the lattice and gives at least kMinTilesPerPhase tiles per
// phase and replica, else the first that applies.  (Measured on L = 48, 32,
// 24 and 16 lattices: 6x6x4 and 8x8x4 are fastest where they give enough
// tiles; with fewer, 4x4x4 keeps more workgroups in flight.)
constexpr int kTileOrder[] = {2, 5, 1, 4, 3};
constexpr int kMinTilesPerPhase = 8;

// Whether tile kind `kind` applies to the tiled sweep of an L lattice: each
// L_a divisible by twice the tile edge (checkerboard of tiles) and the tile
// + halo within the LDS limit.
bool tile_applies_to(const int L[3], int kind, size_t real_size, size_t max_lds) {
    if (kind < 1 || kind > pyro::kPyroTileKinds) return false;
    int d[3];
    pyro::pyro_tile_dims(kind, d);
    for (int a = 0; a < 3; ++a)
        if (L[a] % (2 * d[a]) != 0) return false;
    return pyro::pyro_tile_lds(d[0], d[1], d[2], int(real_size), true) <= max_lds;
}
// The tiled sweep's automatic tile kind (0: none applies).
int auto_tile_kind(const int L[3], size_t real_size, size_t max_lds) {
    for (int kind : kTileOrder) {
        int d[3];
        pyro::pyro_tile_dims(kind, d);
        const int64_t tiles = int64_t(L[0] / (2 * d[0])) * (L[1] / (2 * d[1])) * (L[2] / (2 * d[2]));
        if (tile_applies_to(L, kind, real_size, max_lds) && tiles >= kMinTilesPerPhase) return kind;
    }
    for (int kind : kTileOrder)
        if (tile_applies_to(L, kind, real_size, max_lds)) return kind;
    return 0;
}
// Tile shape of the tiled energy kernel: the sweep's when it has one, else
// the first shape (same order) whose edges divide L (the energy only reads,
// so no checkerboard of tiles is needed) and whose tile + halo fits the LDS
// limit; 0 = the global-memory stencil kernel.
int energy_tile_kind_for(const int L[3], size_t real_size, size_t max_lds, int sweep_tile_kind) {
    if (sweep_tile_kind) return sweep_tile_kind;
    for (int kind : kTileOrder) {
        int d[3];
        pyro::pyro_tile_dims(kind, d);
        if (L[0] % d[0] == 0 && L[1] % d[1] == 0 && L[2] % d[2] == 0 &&
            pyro::pyro_tile_lds(d[0], d[1], d[2], int(real_size), false) <= max_lds)
            return kind;
    }
    return 0;
}
// Per-workgroup LDS limit of the current device.
size_t device_max_lds(int device) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        return size_t(lds);
    return 65536;
}

struct Status : std::runtime_error {
    int code;
    Status(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Status(PYRO_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) hip_check((x), #x)

void require(bool ok, int code, const std::string& msg) {
    if (!ok) throw Status(code, msg);
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return PYRO_OK;
    } catch (const Status& s) {
        g_last_error = s.what();
        return s.code;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return PYRO_E_INVALID;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return PYRO_E_HIP;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PYRO_E_HIP;
    }
}

// Owning device allocation.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        release();
        if (b == 0) b = 8;
        HIPCHK(hipMalloc(&p, b));
        bytes = b;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Captures what body enqueues on s, and on the streams it forks from s and
// joins back, as one instantiated graph.
template <typename F>
hipGraphExec_t capture_graph(hipStream_t s, F&& body) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        if (hipStreamEndCapture(s, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIPCHK(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(ie);
    return exec;
}

// Scratch of the allocation-free entry point (pyro_energy_per_cell_noalloc):
// the energy partials and the field, in a file-scope device array that every
// GPU holds until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: kernels get the address hipGetSymbolAddress
// returns for the current device (noalloc_scratch), never the name.
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 19;  // 4 MiB
__device__ __align__(256) double scratch[kNoAllocScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
size_t g_scratch_lds[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
double* noalloc_scratch(int device) {
    require(device >= 0 && device < kMaxDevices, PYRO_E_UNSUPPORTED, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, PYRO_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device] = static_cast<double*>(p);
        g_scratch_lds[device] = device_max_lds(device);
    }
    return g_scratch_ptr[device];
}

constexpr int kFamilies = 7;
enum Family { kFamSweep = 0, kFamEnergy, kFamExchange, kFamDft, kFamContract, kFamMunu, kFamMunuFlush };

}  // namespace

struct pyro_ctx {
    pyro::LatticeGeom geom;
    std::vector<pyro::CouplingDef> couplings;
    std::set<std::string> names;
    pyro::Topology topo;
    bool is_setup = false;     // device state allocated
    bool topo_ready = false;   // host topology built (needs no device)

    int R = 0, prec = 32, NS = 0, device = 0;
    uint64_t seed = 0;
    double T_ref = 1.0;
    hipStream_t stream = nullptr;

    DevMem S, dir, nbr, cpl, T_slot, slot_of, rep_of, B, acc, cls, counters;
    DevMem acc_wave;          // [R][acc_waves] per-wave accept counters
    int acc_waves = 0;        // wave slots per replica (largest per-colour grid)
    bool count_acc = false;   // sweeps count accepted updates (pyro_sweep with a buffer)
    // XCD-resident sweeps (strategy 6): per-replica stage counters, the
    // kernel's error flag, its wait bound in wall-clock ticks, and whether
    // such a launch has run on this context (then reads check the flag)
    DevMem xcd_bar, xcd_err;
    unsigned long long xcd_ticks = 0;
    bool xcd_used = false;
    DevMem epart, e, Esum, E2sum, nE, att, accp;
    DevMem S_save, dir_save, fields;
    std::vector<int32_t> h_cls;          // all class descriptors, colour-major
    std::vector<int32_t> h_colour_begin; // [n_colours + 1]
    DevMem colour_begin;

    std::vector<double> h_T;             // slot temperatures
    std::vector<double> h_B;             // [R][3]
    uint64_t pass_count = 0;
    uint64_t round_count = 0;

    bool pt_ready = false;
    uint32_t swap_k0 = 0, swap_k1 = 0;

    // pyrochlore J1-J2-J3 couplings defined through pyro_define_J1J2J3 alone:
    // the compile-time-stencil kernels apply
    bool spec = false;
    double sJ1 = 0, sJ2 = 0, sJ3 = 0, sDelta = 1;
    int spec_shells = 0;
    // per-replica (J1, J2, J3, Jx) of a batch of models in one context
    DevMem Jrep;
    bool has_Jrep = false, xxz_rep = false;
    // device anneal schedule: sched[n_sched], energy history [R][n_sched]
    DevMem sched, ehist;
    std::vector<double> h_sched;
    int n_sched = 0, anneal_step = 0;
    hipGraphExec_t agexec = nullptr;
    int a_kind = -1, a_nor = -1, a_nsw = -1, a_ssf = -1, a_strategy = -1;
    unsigned long long graph_gen = 0;  // counts invalidate_graph() calls: captured work is stale when it moves

    // requested: 0 auto, 1 generic per-colour, 2 generic resident,
    // 3 pyrochlore per-colour, 4 pyrochlore resident, 5 pyrochlore tiled
    int strategy = 0;
    int tile_kind = 0;          // tiled strategy tile shape (0: not applicable)
    int exchange_every = 1;     // tempering rounds per replica exchange
    bool fused_passes = false;  // tiled sweeps: a site's passes back to back (pyro_set_fused_passes)
    int tile_forced = 0;        // tile shape requested by pyro_set_tile_shape (0: automatic)
    size_t max_lds = 65536;     // per-workgroup LDS limit of the device
    int tuned_key = -1;        // (kind, n_or) the auto choice was made for
    int tuned_strategy = 1;

    // SSF
    bool ssf_ready = false;
    int n_corr = 0, n_slots = 0, munu_B = 0, munu_T = 0, ssf_chunk = 1;
    bool store2 = true, munu = false;
    std::vector<int32_t> h_alpha, h_beta;
    DevMem F, G, tw0, tw1, tw2, phase, d_alpha, d_beta, sacc, sacc2, sn, ssf_slot;
    DevMem mbuf, mcnt, macc;
    DevMem cbuf;               // buffered contracted correlators (FP32 with munu): [slot][b][c][n_kh]
    // The DFT keeps k2 < ssf_H of every (k0, k1) line (n_kh values per
    // field); the other K follow from -K (the fields are real).  The
    // contraction also accumulates the K outside the kept half with a
    // Nyquist index on axis 0 or 1 (nyquist_01); reads fill in the rest.
    int ssf_H = 0;
    int64_t n_kh = 0;
    DevMem ssf_map;            // tempering slot -> accumulator slot (-1 skip)
    bool has_ssf_map = false;
    bool dft_ready = false;
    int munu_pending = 0;      // samples staged since the last flush (PT rounds)
    std::vector<int> host_munu_count;

    // graph of one tempering round
    hipGraphExec_t gexec = nullptr;
    int g_kind = -1, g_nor = -1, g_nsw = -1, g_ssf = -1, g_strategy = -1;

    // timing
    bool prof = false;
    double prof_ms[kFamilies] = {};
    uint64_t prof_n[kFamilies] = {};
    struct Pending { int fam; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    hipEvent_t marks[2] = {nullptr, nullptr};
    ~pyro_ctx() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (agexec) (void)hipGraphExecDestroy(agexec);

        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto ev : event_pool) (void)hipEventDestroy(ev);
        for (auto ev : marks)
            if (ev) (void)hipEventDestroy(ev);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int64_t N() const { return geom.n_sites(); }
    int64_t nc() const { return geom.n_cells(); }
    // K = (k0 L1 + k1) L2 + k2: its kept index (-1 outside the kept half),
    // its partner -K, and whether its correlators are accumulated directly
    int64_t kept_index(int64_t K) const {
        const int L2 = geom.L[2];
        const int64_t l = K / L2, k2 = K - l * L2;
        return k2 < ssf_H ? l * ssf_H + k2 : -1;
    }
    int64_t neg_k(int64_t K) const {
        const auto& L = geom.L;
        const int64_t k0 = K / (int64_t(L[1]) * L[2]), k1 = (K / L[2]) % L[1], k2 = K % L[2];
        return (((L[0] - k0) % L[0]) * L[1] + (L[1] - k1) % L[1]) * L[2] + (L[2] - k2) % L[2];
    }
    bool nyquist_01(int64_t K) const {
        const auto& L = geom.L;
        const int64_t k0 = K / (int64_t(L[1]) * L[2]), k1 = (K / L[2]) % L[1];
        return (L[0] % 2 == 0 && 2 * k0 == L[0]) || (L[1] % 2 == 0 && 2 * k1 == L[1]);
    }
    size_t real_size() const { return prec == 64 ? 8 : 4; }

    // Copies ordered on the context stream (kernels queued earlier finish
    // first); synchronous so host buffers may be temporaries.
    void h2d(void* dst, const void* src, size_t n) {
        HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    void d2h(void* dst, const void* src, size_t n) {
        HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        check_xcd();
    }
    // Results read after XCD-resident sweeps are valid only if none of their
    // waits timed out and every workgroup ran on its XCD.
    void check_xcd() {
        if (!xcd_used) return;  // set for good once strategy 6 has run (graph replays included)
        int flag = 0;
        HIPCHK(hipMemcpyAsync(&flag, xcd_err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (flag == 0) return;
        HIPCHK(hipMemsetAsync(xcd_err.p, 0, sizeof(int), stream));
        HIPCHK(hipStreamSynchronize(stream));
        throw Status(PYRO_E_HIP, "XCD-resident sweep: a stage wait timed out; results are invalid");
    }

    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    // Run `launch` bracketed by timing events when profiling.
    template <typename L>
    void timed(int fam, L&& launch) {
        if (!prof) {
            HIPCHK(launch());
            return;
        }
        hipEvent_t a = get_event(), b = get_event();
        HIPCHK(hipEventRecord(a, stream));
        HIPCHK(launch());
        HIPCHK(hipEventRecord(b, stream));
        pending.push_back({fam, a, b});
    }
    void collect_timing() {
        if (pending.empty()) return;
        HIPCHK(hipStreamSynchronize(stream));
        for (auto& p : pending) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
            prof_ms[p.fam] += ms;
            prof_n[p.fam] += 1;
            event_pool.push_back(p.a);
            event_pool.push_back(p.b);
        }
        pending.clear();
    }

    template <typename Real>
    pyro::DeviceState<Real> state(const unsigned long long* pass_base) const {
        pyro::DeviceState<Real> st;
        st.S = S.as<Real>();
        st.dir = dir.as<int8_t>();
        st.nbr = nbr.as<int32_t>();
        st.cpl = cpl.as<Real>();
        st.T_slot = T_slot.as<double>();
        st.slot_of = slot_of.as<int32_t>();
        st.B = B.as<Real>();
        st.acc = count_acc ? acc.as<unsigned long long>() : nullptr;
        st.acc_wave = count_acc ? acc_wave.as<unsigned int>() : nullptr;
        st.acc_waves = acc_waves;
        st.pass_base = pass_base;
        st.k0 = uint32_t(seed) ^ pyro::kTagSweep;
        st.k1 = uint32_t(seed >> 32);
        st.T_ref = T_ref;
        st.N = int(N());
        st.n_cells = int(nc());
        st.R = R;
        st.NS = NS;
        st.L0 = geom.L[0];
        st.L1 = geom.L[1];
        st.L2 = geom.L[2];
        st.p0 = topo.period[0];
        st.p1 = topo.period[1];
        st.p2 = topo.period[2];
        st.iso = topo.isotropic;
        return st;
    }
    template <typename Real>
    pyro::PyroCouplings<Real> pyro_couplings() const {
        pyro::PyroCouplings<Real> J;
        J.J1 = Real(sJ1);
        J.J2 = Real(sJ2);
        J.J3 = Real(sJ3);
        J.Jx = Real(sJ1 * (sDelta - 1.0) / 3.0);
        J.shells = spec_shells;
        J.xxz = sDelta != 1.0;
        if (has_Jrep) {
            // the replicas' own couplings replace the context's, Delta included
            J.per_replica = Jrep.as<Real>();
            J.xxz = xxz_rep;
        }
        return J;
    }
    bool resident_fits() const { return nc() <= pyro::kPyroResidentMaxCells; }

    // Tile shape for the tiled strategy (tile_applies_to, auto_tile_kind).
    bool tile_applies(int kind) const { return spec && tile_applies_to(geom.L.data(), kind, real_size(), max_lds); }
    int pick_tile_kind() const {
        if (tile_forced) return tile_applies(tile_forced) ? tile_forced : 0;
        return spec ? auto_tile_kind(geom.L.data(), real_size(), max_lds) : 0;
    }
    // Tile shape of the tiled energy kernel (energy_tile_kind_for).
    int energy_tile_kind() const {
        if (tile_kind) return tile_kind;
        if (!spec) return 0;
        return energy_tile_kind_for(geom.L.data(), real_size(), max_lds, 0);
    }

    template <typename Real>
    pyro::SsfState<Real> ssf_state() const {
        pyro::SsfState<Real> st;
        st.S = S.as<Real>();
        st.F = F.as<Real>();
        st.G = G.as<Real>();
        st.tw0 = tw0.as<Real>();
        st.tw1 = tw1.as<Real>();
        st.tw2 = tw2.as<Real>();
        st.phase = phase.as<Real>();
        st.R = R;
        st.n_sl = geom.n_sl;
        st.n_cells = int(nc());
        st.L0 = geom.L[0];
        st.L1 = geom.L[1];
        st.L2 = geom.L[2];
        st.H = ssf_H;
        st.n_kh = int(n_kh);
        return st;
    }

    void need_setup() const { require(is_setup, PYRO_E_STATE, "setup_lattice() has not been called"); }
    void need_topo() const { require(topo_ready, PYRO_E_STATE, "setup_lattice() has not been called"); }
    void check_replica(int r, bool allow_all) const {
        require((allow_all && r == -1) || (r >= 0 && r < R), PYRO_E_INVALID,
                "replica index out of range");
    }

    // ------------------------------------------------------------------
    void upload_temperatures() {
        h2d(T_slot.p, h_T.data(), sizeof(double) * R);
    }
    void upload_field() {
        if (prec == 64) {
            h2d(B.p, h_B.data(), sizeof(double) * 3 * R);
        } else {
            std::vector<float> f(h_B.begin(), h_B.end());
            h2d(B.p, f.data(), sizeof(float) * 3 * R);
        }
    }
    void identity_permutation() {
        std::vector<int32_t> id(R);
        for (int r = 0; r < R; ++r) id[r] = r;
        h2d(slot_of.p, id.data(), sizeof(int32_t) * R);
        h2d(rep_of.p, id.data(), sizeof(int32_t) * R);
    }

    // One pass of `kind` over all colours, one launch per colour
    // (strategy 1 generic, 3 pyrochlore stencil).
    // epart (stencil kernels only): the pass also leaves each replica's
    // energy partials there (pyro::launch_pyro_colour).
    // fused and ucomp (stencil kernels only): the fused pass order's launch
    // of the first kind with the first OR update, and the uniform component
    // of an OR pass of that order running on its own.
    template <typename Real>
    void pass_colours(int kind, const unsigned long long* pass_base, unsigned long long pass_index,
                      bool use_spec, double* epart = nullptr, bool fused = false, int ucomp = 0) {
        const auto st = state<Real>(pass_base);
        if (use_spec) {
            const auto J = pyro_couplings<Real>();
            for (int c = 0; c < topo.n_colours; ++c)
                timed(kFamSweep, [&] {
                    return pyro::launch_pyro_colour<Real>(st, J, c, kind, pass_index, stream, epart, fused, ucomp);
                });
            return;
        }
        for (int c = 0; c < topo.n_colours; ++c) {
            pyro::ColourList col;
            col.cls = cls.as<int32_t>() + h_colour_begin[c];
            col.n_cls = h_colour_begin[c + 1] - h_colour_begin[c];
            timed(kFamSweep, [&] { return pyro::launch_sweep_colour<Real>(st, col, kind, pass_index, stream); });
        }
    }
    // n_sweeps sweeps of (kind + n_or OR passes) with the given strategy.
    // epart_last (strategies 3 and 4): the last pass also leaves the energy
    // partials of the final state there, the same bits from either.
    template <typename Real>
    void sweeps(int strat, int kind, int n_or, int n_sweeps, const unsigned long long* pass_base,
                unsigned long long pass_index, double* epart_last = nullptr) {
        if (strat == 2) {
            const auto st = state<Real>(pass_base);
            timed(kFamSweep, [&] {
                return pyro::launch_sweep_resident<Real>(st, cls.as<int32_t>(), colour_begin.as<int32_t>(),
                                                         topo.n_colours, kind, n_or, n_sweeps,
                                                         pass_index, stream);
            });
            return;
        }
        if (strat == 5) {
            const auto st = state<Real>(pass_base);
            const auto J = pyro_couplings<Real>();
            unsigned long long p = pass_index;
            for (int sw = 0; sw < n_sweeps; ++sw, p += 1 + n_or)
                for (int phase = 0; phase < 8; ++phase)
                    timed(kFamSweep, [&] {
                        return pyro::launch_pyro_tiled<Real>(st, J, tile_kind, phase, kind, n_or, p, stream,
                                                             fused_order(5, n_or));
                    });
            return;
        }
        if (strat == 4) {
            const auto st = state<Real>(pass_base);
            const auto J = pyro_couplings<Real>();
            timed(kFamSweep, [&] {
                return pyro::launch_pyro_resident<Real>(st, J, kind, n_or, n_sweeps, pass_index, stream,
                                                        epart_last, fused_order(4, n_or));
            });
            return;
        }
        if (strat == 6) {
            const auto st = state<Real>(pass_base);
            const auto J = pyro_couplings<Real>();
            xcd_used = true;
            timed(kFamSweep, [&] {
                const hipError_t e = pyro::launch_pyro_xcd<Real>(st, J, kind, n_or, n_sweeps, pass_index, stream,
                                                                 epart_last, fused_order(6, n_or),
                                                                 xcd_bar.as<unsigned int>(), xcd_err.as<int>(),
                                                                 xcd_ticks);
                require(e != hipErrorNotSupported, PYRO_E_UNSUPPORTED,
                        "the XCD-resident strategy needs its whole grid resident on the device at once");
                return e;
            });
            return;
        }
        unsigned long long p = pass_index;
        if (strat == 3 && fused_order(strat, n_or)) {
            // the fused pass order: per colour the first kind and the first
            // OR update of each site in one launch, then OR passes 2..n_or;
            // OR pass k draws uniform j = kUniformsOf[kind] + k - 1 of the
            // sweep, component j % 5 of block p + j / 5
            const int n_first = pyro::kUniformsOf[kind];
            for (int s = 0; s < n_sweeps; ++s, p += 1 + n_or) {
                const bool last = s == n_sweeps - 1;
                pass_colours<Real>(kind, pass_base, p, true, last && n_or == 1 ? epart_last : nullptr, true);
                for (int k = 2; k <= n_or; ++k) {
                    const int j = n_first + k - 1;
                    pass_colours<Real>(3, pass_base, p + unsigned(j / 5), true,
                                       last && k == n_or ? epart_last : nullptr, false, j % 5);
                }
            }
            return;
        }
        for (int s = 0; s < n_sweeps; ++s) {
            const bool last = s == n_sweeps - 1 && strat == 3;
            pass_colours<Real>(kind, pass_base, p++, strat == 3, last && n_or == 0 ? epart_last : nullptr);
            for (int o = 0; o < n_or; ++o)
                pass_colours<Real>(3, pass_base, p++, strat == 3, last && o == n_or - 1 ? epart_last : nullptr);
        }
    }
    // Whether sweeps of strategy strat with n_or OR passes run in the fused
    // pass order (pyro_set_fused_passes): the stencil strategies 3, 4 and 5,
    // with at least one OR pass to pair with the first kind.
    bool fused_order(int strat, int n_or) const { return fused_passes && n_or >= 1 && strat >= 3; }
    void run_sweeps(int strat, int kind, int n_or, int n_sweeps, const unsigned long long* pass_base,
                    unsigned long long pass_index, double* epart_last = nullptr) {
        if (prec == 64) sweeps<double>(strat, kind, n_or, n_sweeps, pass_base, pass_index, epart_last);
        else sweeps<float>(strat, kind, n_or, n_sweeps, pass_base, pass_index, epart_last);
    }
    // Whether run_sweeps(strat, ...) with n_sweeps > 0 can leave the energy
    // partials (energy_finish then replaces energy): the stencil strategies
    // 3 and 4, which leave identical partials (so the autotuner's choice
    // between them never changes a sampled energy), with the observables on
    // the stencil too.
    bool sweeps_leave_energy(int strat) const {
        return (strat == 3 || strat == 4 || strat == 6) && stencil_observables();
    }

    // Pick the faster launch strategy for (kind, n_or) by timing both on a
    // saved copy of the state, then restore it (results do not depend on the
    // strategy: both use the same per-site counters and colour order).
    int choose_strategy(int kind, int n_or) {
        if (strategy != 0) {
            require(strategy <= 2 || spec, PYRO_E_UNSUPPORTED,
                    "strategies 3 to 6 need couplings defined by define_J1J2J3 alone");
            require(strategy >= 3 || !has_Jrep, PYRO_E_UNSUPPORTED,
                    "per-replica couplings need the pyrochlore strategies 3 to 6");
            require(strategy != 4 || resident_fits(), PYRO_E_UNSUPPORTED,
                    "the pyrochlore resident strategy needs at most 2048 cells");
            require(strategy != 5 || tile_kind > 0, PYRO_E_UNSUPPORTED,
                    "the tiled strategy needs each L divisible by twice the tile edge (8 or 12)");
            return strategy;
        }
        // Large lattices take the tiled schedule (a fixed rule, so results do
        // not depend on timing).
        if (tile_kind > 0 && nc() >= 4096) {
            tuned_key = kind * 64 + n_or;
            tuned_strategy = 5;
            return 5;
        }
        const int key = kind * 64 + n_or;
        if (tuned_key == key) return tuned_strategy;
        std::vector<int> cand;
        if (spec) {
            cand.push_back(3);
            if (resident_fits()) cand.push_back(4);
            cand.push_back(6);  // dropped below if its grid cannot be resident at once
        } else {
            cand = {1, 2};
        }
        if (cand.size() == 1) {
            tuned_key = key;
            tuned_strategy = cand[0];
            return cand[0];
        }
        const size_t sbytes = size_t(R) * 3 * N() * real_size();
        const size_t dbytes = size_t(R) * N();
        if (!S_save.p) {
            S_save.alloc(sbytes);
            dir_save.alloc(dbytes);
        }
        HIPCHK(hipMemcpyAsync(S_save.p, S.p, sbytes, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(dir_save.p, dir.p, dbytes, hipMemcpyDeviceToDevice, stream));
        hipEvent_t a = get_event(), b = get_event();
        const bool was_prof = prof;
        prof = false;
        float best = 1e30f;
        int best_s = spec ? 3 : 1;
        for (int s : cand) {
            if (s == 6) {
                // the XCD-resident kernel applies when its whole grid fits on
                // the device at once and its waits complete
                try {
                    run_sweeps(s, kind, n_or, 1, nullptr, pass_count);
                    HIPCHK(hipStreamSynchronize(stream));
                    check_xcd();
                } catch (const Status&) {
                    (void)hipGetLastError();
                    continue;
                }
            } else {
                run_sweeps(s, kind, n_or, 1, nullptr, pass_count);  // warm-up
            }
            HIPCHK(hipEventRecord(a, stream));
            run_sweeps(s, kind, n_or, 3, nullptr, pass_count);
            HIPCHK(hipEventRecord(b, stream));
            HIPCHK(hipEventSynchronize(b));
            if (s == 6) {
                try {
                    check_xcd();
                } catch (const Status&) {
                    continue;
                }
            }
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, a, b));
            if (ms < best) {
                best = ms;
                best_s = s;
            }
        }
        prof = was_prof;
        event_pool.push_back(a);
        event_pool.push_back(b);
        HIPCHK(hipMemcpyAsync(S.p, S_save.p, sbytes, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(dir.p, dir_save.p, dbytes, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        tuned_key = key;
        tuned_strategy = best_s;
        return best_s;
    }

    // Per-replica energies (and with accumulate the slot sums) from nblk
    // partials per replica in epart.
    void energy_finish(bool accumulate, int nblk) {
        double* es = accumulate ? Esum.as<double>() : nullptr;
        double* e2 = accumulate ? E2sum.as<double>() : nullptr;
        unsigned long long* ne = accumulate ? nE.as<unsigned long long>() : nullptr;
        timed(kFamEnergy, [&] {
            return pyro::launch_energy_finish(epart.as<double>(), nblk, int(nc()), e.as<double>(),
                                              slot_of.as<int32_t>(), es, e2, ne, R, stream);
        });
    }

    void energy(bool accumulate) {
        double* es = accumulate ? Esum.as<double>() : nullptr;
        double* e2 = accumulate ? E2sum.as<double>() : nullptr;
        unsigned long long* ne = accumulate ? nE.as<unsigned long long>() : nullptr;
        if (stencil_observables()) {
            // lattices a tile shape divides: tiles read once into LDS
            // (partials per tile); others: one 64-site block per partial
            int nblk = pyro::pyro_energy_blocks(int(N()));
            const int ek = energy_tile_kind();
            timed(kFamEnergy, [&] {
                if (ek != 0)
                    return prec == 64 ? pyro::launch_pyro_energy_tiled<double>(state<double>(nullptr),
                                                                               pyro_couplings<double>(), ek,
                                                                               epart.as<double>(), &nblk, stream)
                                      : pyro::launch_pyro_energy_tiled<float>(state<float>(nullptr),
                                                                              pyro_couplings<float>(), ek,
                                                                              epart.as<double>(), &nblk, stream);
                return prec == 64
                           ? pyro::launch_pyro_energy<double>(state<double>(nullptr), pyro_couplings<double>(),
                                                              epart.as<double>(), stream)
                           : pyro::launch_pyro_energy<float>(state<float>(nullptr), pyro_couplings<float>(),
                                                             epart.as<double>(), stream);
            });
            energy_finish(accumulate, nblk);
            return;
        }
        timed(kFamEnergy, [&] {
            return prec == 64 ? pyro::launch_energy<double>(state<double>(nullptr), epart.as<double>(),
                                                            e.as<double>(), es, e2, ne, stream)
                              : pyro::launch_energy<float>(state<float>(nullptr), epart.as<double>(),
                                                           e.as<double>(), es, e2, ne, stream);
        });
    }

    // SSF sample of replicas [0, R) using the device slot array `slots`.
    // SSF sample on stream s (kernel timing records on the context stream,
    // so s is the context stream whenever timing is on).
    // One SSF sample of every replica into accumulator slots[r].
    void ssf_sample_device(const int32_t* slots, hipStream_t s) {
        for (int r0 = 0; r0 < R; r0 += ssf_chunk) {
            const int rc = std::min(ssf_chunk, R - r0);
            pyro::CorrArgs c;
            c.slot_of = slots;
            c.alpha = d_alpha.as<int32_t>();
            c.beta = d_beta.as<int32_t>();
            c.acc = sacc.as<double>();
            c.acc2 = store2 ? sacc2.as<double>() : nullptr;
            c.n_samples = sn.as<unsigned long long>();
            c.n_corr = n_corr;
            if (prec == 64) {
                const auto st = ssf_state<double>();
                timed(kFamDft, [&] { return pyro::launch_dft<double>(st, r0, rc, s); });
                timed(kFamContract, [&] { return pyro::launch_contract<double>(st, c, r0, rc, s); });
                if (munu) timed(kFamMunu, [&] { return pyro::launch_munu_stage<double>(st, slots, r0, rc, munu_args(), s); });
            } else {
                const auto st = ssf_state<float>();
                timed(kFamDft, [&] { return pyro::launch_dft<float>(st, r0, rc, s); });
                if (munu)  // contraction and staging from one read of the fields
                    timed(kFamContract, [&] { return pyro::launch_contract_stage(st, c, r0, rc, munu_args(), s); });
                else
                    timed(kFamContract, [&] { return pyro::launch_contract<float>(st, c, r0, rc, s); });
            }
        }
    }
    pyro::MunuArgs munu_args() const {
        pyro::MunuArgs m;
        m.buf = mbuf.as<float>();
        m.cnt = mcnt.as<unsigned int>();
        m.acc = macc.as<double>();
        m.Bmax = munu_B;
        m.n_k = int(n_kh);
        m.n_slots = n_slots;
        m.n_f = 3 * geom.n_sl;
        m.T = munu_T;
        if (cbuf.p) {
            m.cbuf = cbuf.as<float2>();
            m.acc_s = sacc.as<double>();
            m.acc_s2 = store2 ? sacc2.as<double>() : nullptr;
            m.n_corr = n_corr;
            m.H = ssf_H;
            m.L2 = geom.L[2];
            m.n_k_full = int(nc());
        }
        return m;
    }
    void munu_flush() {
        if (!munu) return;
        // the buffered contracted samples (timed with the contraction), then
        // the sublattice-resolved products; the latter resets the row counts
        if (cbuf.p) timed(kFamContract, [&] { return pyro::launch_corr_flush(munu_args(), stream); });
        timed(kFamMunuFlush, [&] { return pyro::launch_munu_flush(munu_args(), stream); });
        munu_pending = 0;
        std::fill(host_munu_count.begin(), host_munu_count.end(), 0);
    }

    // DFT work buffers, per-axis twiddles e^{-2 pi i j / L_a} and the
    // sublattice phases e^{-i q(K).r_mu} (q with the centred K').
    void ensure_dft() {
        if (dft_ready) return;
        const size_t rs = real_size();
        const int n_sl = geom.n_sl;
        const auto& L = geom.L;
        ssf_H = pyro::dft_kept_k2(int(rs), L[0], L[1], L[2], 3ll * n_sl * nc());
        n_kh = int64_t(L[0]) * L[1] * ssf_H;
        const size_t per_rep = size_t(3) * n_sl * n_kh * 2 * rs;
        ssf_chunk = int(std::max<size_t>(1, std::min<size_t>(R, (size_t(2) << 30) / per_rep)));
        F.alloc(per_rep * ssf_chunk);
        G.alloc(per_rep * ssf_chunk);
        auto upload_real = [&](DevMem& m, const std::vector<double>& v) {
            m.alloc(v.size() * rs);
            if (rs == 8) {
                h2d(m.p, v.data(), v.size() * 8);
            } else {
                std::vector<float> f(v.begin(), v.end());
                h2d(m.p, f.data(), f.size() * 4);
            }
        };
        DevMem* tws[3] = {&tw0, &tw1, &tw2};
        for (int a = 0; a < 3; ++a) {
            std::vector<double> tw(2 * L[a]);
            for (int j = 0; j < L[a]; ++j) {
                const double th = 2.0 * M_PI * double(j) / double(L[a]);
                tw[2 * j] = std::cos(th);
                tw[2 * j + 1] = -std::sin(th);
            }
            upload_real(*tws[a], tw);
        }
        const int64_t n_k = nc();
        std::vector<double> ph(size_t(2) * n_sl * n_k);
        for (int mu = 0; mu < n_sl; ++mu) {
            const auto& r = geom.sl_pos[mu];
            for (int64_t K = 0; K < n_k; ++K) {
                const int Ks[3] = {int(K / (int64_t(L[1]) * L[2])), int((K / L[2]) % L[1]), int(K % L[2])};
                double arg = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const int Kc = (2 * Ks[a] < L[a]) ? Ks[a] : Ks[a] - L[a];
                    arg += 2.0 * M_PI * double(Kc) / double(geom.cell[a] * L[a]) * double(r[a]);
                }
                ph[2 * (mu * n_k + K)] = std::cos(arg);
                ph[2 * (mu * n_k + K) + 1] = -std::sin(arg);
            }
        }
        upload_real(phase, ph);
        dft_ready = true;
    }

    void invalidate_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        if (agexec) (void)hipGraphExecDestroy(agexec);
        agexec = nullptr;
        ++graph_gen;
    }

    // Anneal steps (pyro_anneal_steps, pyro_batch_anneal_steps): argument
    // checks, the device counters (pass base, round, schedule step) set from
    // the host's, the step's temperatures from the schedule, and after the
    // sweeps the step's SSF sample, energy and energy record.
    void anneal_check(int kind, int n_or, int n_sweep, int sample_ssf, int n_steps) const {
        need_setup();
        require(n_sched > 0, PYRO_E_STATE, "anneal_schedule() has not been called");
        require(kind == PYRO_METROPOLIS || kind == PYRO_LIFTED || kind == PYRO_LIFTED_ROT, PYRO_E_INVALID,
                "sweep kind must be METROPOLIS, LIFTED or LIFTED_ROT");
        require(n_or >= 0 && n_sweep >= 0 && n_steps >= 0, PYRO_E_INVALID, "negative count");
        require(anneal_step + n_steps <= n_sched, PYRO_E_INVALID, "steps beyond the end of the schedule");
        require(!sample_ssf || ssf_ready, PYRO_E_STATE, "ssf_configure() has not been called");
        require(!sample_ssf || has_ssf_map || n_slots >= R, PYRO_E_INVALID,
                "SSF sampling in anneal steps needs a slot map or one accumulator slot per replica");
    }
    void anneal_upload_counters() {
        const unsigned long long hc[3] = {pass_count, round_count, (unsigned long long)anneal_step};
        h2d(counters.p, hc, sizeof(hc));
    }
    void anneal_set_T() {
        HIPCHK(pyro::launch_anneal_set_T(T_slot.as<double>(), R, sched.as<double>(), n_sched,
                                         counters.as<unsigned long long>() + 2, stream));
    }
    void anneal_observe(int sample_ssf, bool fused_energy, unsigned long long ppr) {
        auto* ctr = counters.as<unsigned long long>();
        if (sample_ssf) {
            if (has_ssf_map) {  // replica r (slot r) -> accumulator map[r]
                HIPCHK(pyro::launch_map_slots(slot_of.as<int32_t>(), ssf_map.as<int32_t>(), ssf_slot.as<int32_t>(),
                                              R, stream));
                ssf_sample_device(ssf_slot.as<int32_t>(), stream);
            } else {
                ssf_sample_device(slot_of.as<int32_t>(), stream);
            }
        }
        if (fused_energy) energy_finish(false, pyro::pyro_colour_energy_blocks(int(nc())));
        else energy(false);
        HIPCHK(pyro::launch_anneal_record(e.as<double>(), R, ehist.as<double>(), n_sched, ctr + 2, ctr, ppr, stream));
    }
    // Host bookkeeping after one step is enqueued; true when the mu-nu
    // buffer is full and must be flushed before the next step.
    bool anneal_advance(int sample_ssf, unsigned long long ppr) {
        pass_count += ppr;
        anneal_step += 1;
        if (sample_ssf && munu) {
            munu_pending += 1;
            return munu_pending >= munu_B;
        }
        return false;
    }
    // Stencil kernels for field and energy: pyrochlore couplings, unless a
    // table strategy (1, 2) is forced without per-replica couplings.
    bool stencil_observables() const { return spec && (has_Jrep || strategy == 0 || strategy >= 3); }
};

// A batch of contexts stepped together (pyro_batch_anneal_steps): one graph
// per step for all of them, in which the members that take the per-colour
// stencil strategies sweep in one persistent launch (pyro::launch_pyro_batch)
// and the others (the tiled strategy on large lattices, forced strategies)
// run their own sweeps in parallel branches.  Each member's work runs on its
// own stream, forked from the batch's stream and joined back.
struct pyro_batch {
    std::vector<pyro_ctx*> ctx;
    hipStream_t stream = nullptr;
    // fork/join events: fork, the persistent launch's end, the batch's end,
    // then per member its pre-sweep end and its step end
    std::vector<hipEvent_t> ev;
    DevMem members, work, bar;
    std::vector<unsigned char> h_members;  // the uploaded member table, bytes
    std::vector<int32_t> h_work;           // the uploaded work table
    std::vector<char> in_launch;           // member sweeps in the persistent launch
    int per_xcd = 0, n_items = 0, blocks_per_worker = 0, launch_members = 0;
    hipGraphExec_t gexec = nullptr;
    std::vector<long long> g_key;          // what gexec was captured for
    ~pyro_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Whether member c sweeps in the batch's persistent launch: a context of
// the per-colour stencil strategies left to the automatic choice (not the
// tiled strategy's large lattices, whose fixed rule gives a different
// visiting order), or one forced to the XCD-resident sweep (whose results
// the launch reproduces, and which beside it would be a second persistent
// kernel competing for residency), with replicas below 2^24 sites (the
// launch's 24-bit byte offsets).
bool batch_launch_eligible(const pyro_ctx* c) {
    const bool autom = c->strategy == 0 && !(c->tile_kind > 0 && c->nc() >= 4096);
    return c->spec && (autom || c->strategy == 6) && c->N() < (int64_t(1) << 24);
}

// Places every replica of the launch's members whole on one XCD: worker
// counts ceil(bpr / k) for the fewest blocks per worker k that fit, the
// replicas (largest first) onto the least loaded XCD.  Fills the work
// table; false when not even one worker per replica fits.
bool batch_plan(const std::vector<pyro_ctx*>& m, int capacity, std::vector<int32_t>& work, int& per_xcd, int& k_out) {
    struct Item { int member, r, bpr; };
    std::vector<Item> items;
    int max_bpr = 1;
    for (int i = 0; i < int(m.size()); ++i) {
        const int bpr = int((m[i]->nc() + 63) / 64) * (m[i]->spec_shells == 3 ? 1 : 4);
        max_bpr = std::max(max_bpr, bpr);
        for (int r = 0; r < m[i]->R; ++r) items.push_back({i, r, bpr});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.bpr > b.bpr; });
    for (int k = 1; k <= max_bpr; ++k) {
        std::vector<std::vector<int32_t>> xcd(8);
        std::vector<long long> load(8, 0);
        for (const Item& it : items) {
            const int nw = (it.bpr + k - 1) / k;
            const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
            load[x] += nw;
            for (int w = 0; w < nw; ++w) xcd[x].insert(xcd[x].end(), {it.member, it.r, w, nw});
        }
        const long long px = *std::max_element(load.begin(), load.end());
        if (8 * px > capacity) continue;
        per_xcd = int(px);
        k_out = k;
        work.assign(size_t(8) * size_t(px) * 4, -1);
        for (int x = 0; x < 8; ++x) std::copy(xcd[x].begin(), xcd[x].end(), work.begin() + size_t(x) * px * 4);
        return true;
    }
    return false;
}

template <typename Real>
void batch_members(pyro_batch* b, const std::vector<pyro_ctx*>& m, std::vector<unsigned char>& bytes) {
    std::vector<pyro::BatchMember<Real>> t(m.size());
    int item = 0;
    for (size_t i = 0; i < m.size(); ++i) {
        t[i].st = m[i]->state<Real>(m[i]->counters.as<unsigned long long>());
        t[i].J = m[i]->pyro_couplings<Real>();
        t[i].epart = m[i]->epart.as<double>();
        t[i].err = m[i]->xcd_err.as<int>();
        t[i].timeout = m[i]->xcd_ticks;
        t[i].pass_index = 0;
        t[i].item0 = item;
        item += m[i]->R;
    }
    b->n_items = item;
    const auto* p = reinterpret_cast<const unsigned char*>(t.data());
    bytes.assign(p, p + sizeof(pyro::BatchMember<Real>) * t.size());
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* pyro_last_error(void) { return g_last_error.c_str(); }

int pyro_version(void) { return 10000; }

int pyro_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pyro_create(int n_sl, const int32_t* sl_pos, const int32_t* sl_kind, const int32_t cell[3],
                const int32_t L[3], int n_replicas, int precision, uint64_t seed, pyro_ctx** out) {
    return guarded([&] {
        require(out != nullptr, PYRO_E_INVALID, "out is NULL");
        *out = nullptr;
        require(n_sl > 0 && sl_pos && sl_kind && cell && L, PYRO_E_INVALID, "lattice description incomplete");
        require(n_replicas > 0 && n_replicas <= 65535, PYRO_E_INVALID, "n_replicas must be in [1, 65535]");
        require(precision == PYRO_FP32 || precision == PYRO_FP64, PYRO_E_INVALID, "precision must be 32 or 64");
        auto ctx = std::make_unique<pyro_ctx>();
        ctx->geom.n_sl = n_sl;
        for (int s = 0; s < n_sl; ++s) {
            ctx->geom.sl_pos.push_back({sl_pos[3 * s], sl_pos[3 * s + 1], sl_pos[3 * s + 2]});
            ctx->geom.sl_kind.push_back(sl_kind[s]);
        }
        ctx->geom.cell = {cell[0], cell[1], cell[2]};
        ctx->geom.L = {L[0], L[1], L[2]};
        ctx->geom.validate();
        ctx->R = n_replicas;
        ctx->prec = precision;
        ctx->seed = seed;
        ctx->h_T.assign(n_replicas, 1.0);
        ctx->h_B.assign(size_t(3) * n_replicas, 0.0);
        *out = ctx.release();
    });
}

int pyro_destroy(pyro_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    });
}

int pyro_define_coupling(pyro_ctx* ctx, const char* name, int n_kinds, const int32_t* counts,
                         const int32_t* rel_vecs, const double J[9]) {
    return guarded([&] {
        require(ctx && name && J && (n_kinds == 0 || counts), PYRO_E_INVALID, "NULL argument");
        require(!ctx->topo_ready, PYRO_E_STATE, "couplings must be defined before setup_lattice()");
        require(!ctx->names.count(name), PYRO_E_LOGIC, "Coupling names must be unique");
        pyro::CouplingDef c;
        c.name = name;
        size_t off = 0;
        for (int k = 0; k < n_kinds; ++k) {
            require(counts[k] >= 0, PYRO_E_INVALID, "negative displacement count");
            std::vector<pyro::ivec3> v;
            for (int i = 0; i < counts[k]; ++i, ++off)
                v.push_back({rel_vecs[3 * off], rel_vecs[3 * off + 1], rel_vecs[3 * off + 2]});
            c.rel.push_back(std::move(v));
        }
        for (int e = 0; e < 9; ++e) c.J[e] = J[e];
        ctx->names.insert(c.name);
        ctx->couplings.push_back(std::move(c));
        ctx->spec = false;
    });
}

int pyro_define_J1J2J3(pyro_ctx* ctx, double J1, double J2, double J3, double Delta) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(!ctx->topo_ready, PYRO_E_STATE, "couplings must be defined before setup_lattice()");
        const auto ref = pyro::pyrochlore_cell(1, 1, 1);
        bool pyro_cell = ctx->geom.n_sl == 16 && ctx->geom.cell == ref.cell;
        for (int s = 0; pyro_cell && s < 16; ++s)
            pyro_cell = ctx->geom.sl_pos[s] == ref.sl_pos[s] && ctx->geom.sl_kind[s] == ref.sl_kind[s];
        require(pyro_cell, PYRO_E_INVALID, "define_J1J2J3 needs the 16-site pyrochlore cubic cell");
        const bool first = ctx->couplings.empty();
        for (auto& c : pyro::pyrochlore_J1J2J3(J1, J2, J3, Delta)) {
            require(!ctx->names.count(c.name), PYRO_E_LOGIC, "Coupling names must be unique");
            ctx->names.insert(c.name);
            ctx->couplings.push_back(std::move(c));
        }
        ctx->spec = first && J1 != 0.0;
        ctx->sJ1 = J1;
        ctx->sJ2 = J2;
        ctx->sJ3 = J3;
        ctx->sDelta = Delta;
        ctx->spec_shells = J3 != 0.0 ? 3 : (J2 != 0.0 ? 2 : 1);
    });
}

int pyro_pyrochlore_shell(int which, int32_t counts[4], int32_t* vecs, int cap) {
    return guarded([&] {
        require(which >= 1 && which <= 4 && counts, PYRO_E_INVALID, "which must be 1..4");
        const auto sh = pyro::pyrochlore_shell(which);
        int off = 0;
        for (int p = 0; p < 4; ++p) {
            counts[p] = int32_t(sh[p].size());
            for (const auto& d : sh[p]) {
                if (vecs && off < cap)
                    for (int a = 0; a < 3; ++a) vecs[3 * off + a] = d[a];
                ++off;
            }
        }
    });
}

int pyro_setup_lattice(pyro_ctx* ctx) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(!ctx->is_setup, PYRO_E_STATE, "setup_lattice() already called");
        if (!ctx->topo_ready) {
            ctx->topo = pyro::build_topology(ctx->geom, ctx->couplings);
            ctx->topo_ready = true;
        }
        auto& t = ctx->topo;
        const int ns_list[] = {6, 8, 18, 30, 32};
        ctx->NS = 0;
        for (int v : ns_list)
            if (t.n_slot <= v) {
                ctx->NS = v;
                break;
            }
        require(ctx->NS > 0, PYRO_E_UNSUPPORTED, "more than 32 coupled neighbours per site");
        // the topology above is host-only; everything below needs a device
        int ndev = 0;
        require(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, PYRO_E_HIP, "no HIP device available");
        HIPCHK(hipGetDevice(&ctx->device));
        if (!ctx->stream) HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->max_lds = device_max_lds(ctx->device);
        const int64_t N = ctx->N();
        const int R = ctx->R;
        const int NS = ctx->NS;
        const size_t rs = ctx->real_size();

        // neighbour table, slot-major and padded to NS slots with the site itself
        std::vector<int32_t> nbr(size_t(NS) * N);
        for (int64_t i = 0; i < N; ++i)
            for (int k = 0; k < NS; ++k)
                nbr[size_t(k) * N + i] = k < t.n_slot ? t.nbr[size_t(i) * t.n_slot + k] : int32_t(i);
        ctx->nbr.alloc(nbr.size() * sizeof(int32_t));
        ctx->h2d(ctx->nbr.p, nbr.data(), nbr.size() * sizeof(int32_t));

        // per-slot couplings
        const int n_sl = ctx->geom.n_sl;
        std::vector<double> cp;
        if (t.isotropic) {
            cp.assign(size_t(n_sl) * NS, 0.0);
            for (int s = 0; s < n_sl; ++s)
                for (int k = 0; k < t.n_slot; ++k) cp[size_t(s) * NS + k] = t.cslot[size_t(s) * t.n_slot + k];
        } else {
            cp.assign(size_t(n_sl) * NS * 9, 0.0);
            for (int s = 0; s < n_sl; ++s)
                for (int k = 0; k < t.n_slot; ++k)
                    for (int e = 0; e < 9; ++e)
                        cp[(size_t(s) * NS + k) * 9 + e] = t.Jslot[(size_t(s) * t.n_slot + k) * 9 + e];
        }
        ctx->cpl.alloc(cp.size() * rs);
        if (rs == 8) {
            ctx->h2d(ctx->cpl.p, cp.data(), cp.size() * 8);
        } else {
            std::vector<float> f(cp.begin(), cp.end());
            ctx->h2d(ctx->cpl.p, f.data(), f.size() * 4);
        }

        // colour classes, colour-major
        const int P = t.period[0] * t.period[1] * t.period[2];
        if (ctx->spec) {
            // the stencil kernels colour by pyro_sl (4) or by sublattice (16)
            const int ncol = ctx->spec_shells == 3 ? 16 : 4;
            bool ok = P == 1 && t.n_colours == ncol;
            for (int sl = 0; ok && sl < 16; ++sl) ok = t.class_colour[sl] == (ncol == 16 ? sl : sl % 4);
            ctx->spec = ok;
        }
        ctx->h_cls.clear();
        ctx->h_colour_begin.assign(1, 0);
        for (int c = 0; c < t.n_colours; ++c) {
            for (int cl : t.colour_classes[c]) {
                const int sl = cl / P;
                int res = cl % P;
                const int r2 = res % t.period[2];
                res /= t.period[2];
                const int r1 = res % t.period[1];
                const int r0 = res / t.period[1];
                ctx->h_cls.push_back(sl | (r0 << 16) | (r1 << 20) | (r2 << 24));
            }
            ctx->h_colour_begin.push_back(int32_t(ctx->h_cls.size()));
        }
        ctx->cls.alloc(ctx->h_cls.size() * sizeof(int32_t));
        ctx->h2d(ctx->cls.p, ctx->h_cls.data(), ctx->h_cls.size() * sizeof(int32_t));
        ctx->colour_begin.alloc(ctx->h_colour_begin.size() * sizeof(int32_t));
        ctx->h2d(ctx->colour_begin.p, ctx->h_colour_begin.data(), ctx->h_colour_begin.size() * sizeof(int32_t));

        // state
        ctx->S.alloc(size_t(R) * 3 * N * rs);
        ctx->dir.alloc(size_t(R) * N);
        ctx->T_slot.alloc(sizeof(double) * R);
        ctx->slot_of.alloc(sizeof(int32_t) * R);
        ctx->rep_of.alloc(sizeof(int32_t) * R);
        ctx->B.alloc(rs * 3 * R);
        ctx->acc.alloc(sizeof(unsigned long long) * R);
        {
            // wave slots of the largest per-colour grid: 256-thread blocks of
            // 4 waves; generic kernels spc-site classes, stencil kernels
            // whole sublattices (4 per launch with 4 colours)
            const int64_t nc = ctx->nc();
            const int64_t spc = nc / (ctx->topo.period[0] * ctx->topo.period[1] * ctx->topo.period[2]);
            int64_t blocks = 4 * ((nc + 255) / 256);
            for (int c = 0; c < ctx->topo.n_colours; ++c)
                blocks = std::max<int64_t>(blocks, ((spc + 255) / 256) *
                                                       (ctx->h_colour_begin[c + 1] - ctx->h_colour_begin[c]));
            ctx->acc_waves = int(4 * blocks);
            ctx->acc_wave.alloc(sizeof(unsigned int) * size_t(R) * size_t(ctx->acc_waves));
        }
        ctx->counters.alloc(sizeof(unsigned long long) * 3);  // pass base, round, anneal step
        {
            // XCD-resident sweeps: stage counters, error flag, a 0.25 s bound
            // on every wait in ticks of the device's wall clock
            ctx->xcd_bar.alloc(sizeof(unsigned int) * 32 * (R + 8));  // R stage counters, 8 XCD tickets, 128 B apart
            ctx->xcd_err.alloc(sizeof(int));
            HIPCHK(hipMemset(ctx->xcd_err.p, 0, sizeof(int)));
            int dev = 0, khz = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            ctx->xcd_ticks = (unsigned long long)(khz > 0 ? khz : 100000) * 250ull;
        }
        ctx->epart.alloc(sizeof(double) * R *
                         std::max({pyro::energy_blocks(int(N)), pyro::pyro_energy_blocks(int(N)),
                                   pyro::pyro_colour_energy_blocks(int(ctx->nc()))}));
        ctx->e.alloc(sizeof(double) * R);
        ctx->Esum.alloc(sizeof(double) * R);
        ctx->E2sum.alloc(sizeof(double) * R);
        ctx->nE.alloc(sizeof(unsigned long long) * R);
        ctx->att.alloc(sizeof(unsigned long long) * R);
        ctx->accp.alloc(sizeof(unsigned long long) * R);
        ctx->fields.alloc(size_t(3) * N * rs);
        HIPCHK(hipMemset(ctx->S.p, 0, ctx->S.bytes));
        HIPCHK(hipMemset(ctx->dir.p, 1, ctx->dir.bytes));
        HIPCHK(hipMemset(ctx->counters.p, 0, ctx->counters.bytes));
        HIPCHK(hipMemset(ctx->Esum.p, 0, ctx->Esum.bytes));
        HIPCHK(hipMemset(ctx->E2sum.p, 0, ctx->E2sum.bytes));
        HIPCHK(hipMemset(ctx->nE.p, 0, ctx->nE.bytes));
        HIPCHK(hipMemset(ctx->att.p, 0, ctx->att.bytes));
        HIPCHK(hipMemset(ctx->accp.p, 0, ctx->accp.bytes));
        HIPCHK(hipDeviceSynchronize());
        ctx->identity_permutation();
        ctx->upload_temperatures();
        ctx->upload_field();
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->tile_kind = ctx->pick_tile_kind();
        ctx->is_setup = true;
    });
}

int pyro_info(pyro_ctx* ctx, int64_t info[8]) {
    return guarded([&] {
        require(ctx && info, PYRO_E_INVALID, "NULL argument");
        info[0] = ctx->N();
        info[1] = ctx->nc();
        info[2] = ctx->geom.n_sl;
        info[3] = ctx->R;
        info[4] = ctx->prec;
        info[5] = ctx->topo_ready ? ctx->topo.n_slot : -1;
        info[6] = ctx->topo_ready ? ctx->topo.n_colours : -1;
        info[7] = ctx->topo_ready ? (ctx->topo.isotropic ? 1 : 0) : -1;
    });
}

int pyro_site_positions(pyro_ctx* ctx, int32_t* pos) {
    return guarded([&] {
        require(ctx && pos, PYRO_E_INVALID, "NULL argument");
        for (int64_t i = 0; i < ctx->N(); ++i) {
            const auto p = ctx->geom.position(i);
            pos[3 * i] = p[0];
            pos[3 * i + 1] = p[1];
            pos[3 * i + 2] = p[2];
        }
    });
}

int pyro_neighbours(pyro_ctx* ctx, int32_t* nbr, double* J) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_topo();
        const auto& t = ctx->topo;
        const int64_t N = ctx->N();
        if (nbr) std::copy(t.nbr.begin(), t.nbr.end(), nbr);
        if (J) {
            for (int64_t i = 0; i < N; ++i) {
                const int s = int(i / ctx->nc());
                for (int k = 0; k < t.n_slot; ++k)
                    for (int e = 0; e < 9; ++e)
                        J[(size_t(i) * t.n_slot + k) * 9 + e] = t.Jslot[(size_t(s) * t.n_slot + k) * 9 + e];
            }
        }
    });
}

int pyro_colouring(pyro_ctx* ctx, int32_t* colour_of_site) {
    return guarded([&] {
        require(ctx && colour_of_site, PYRO_E_INVALID, "NULL argument");
        ctx->need_topo();
        const auto& t = ctx->topo;
        const auto& L = ctx->geom.L;
        const int P = t.period[0] * t.period[1] * t.period[2];
        for (int64_t i = 0; i < ctx->N(); ++i) {
            const int s = int(i / ctx->nc());
            int64_t c = i % ctx->nc();
            const int n2 = int(c % L[2]);
            c /= L[2];
            const int n1 = int(c % L[1]);
            const int n0 = int(c / L[1]);
            const int cl = s * P + ((n0 % t.period[0]) * t.period[1] + (n1 % t.period[1])) * t.period[2] +
                           (n2 % t.period[2]);
            colour_of_site[i] = t.class_colour[cl];
        }
    });
}

int pyro_set_field(pyro_ctx* ctx, int replica, const double B[3]) {
    return guarded([&] {
        require(ctx && B, PYRO_E_INVALID, "NULL argument");
        ctx->check_replica(replica, true);
        for (int r = 0; r < ctx->R; ++r)
            if (replica == -1 || r == replica)
                for (int a = 0; a < 3; ++a) ctx->h_B[3 * r + a] = B[a];
        if (ctx->is_setup) {
            ctx->upload_field();
            HIPCHK(hipStreamSynchronize(ctx->stream));
        }
    });
}

int pyro_set_T_ref(pyro_ctx* ctx, double T_ref) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(T_ref > 0 && std::isfinite(T_ref), PYRO_E_INVALID, "T_ref must be positive");
        ctx->T_ref = T_ref;
        ctx->invalidate_graph();
    });
}

int pyro_set_temperatures(pyro_ctx* ctx, const double* T) {
    return guarded([&] {
        require(ctx && T, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        for (int r = 0; r < ctx->R; ++r)
            require(T[r] > 0 && std::isfinite(T[r]), PYRO_E_INVALID, "temperatures must be positive");
        ctx->h_T.assign(T, T + ctx->R);
        ctx->upload_temperatures();
        ctx->identity_permutation();
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_set_spins(pyro_ctx* ctx, int replica, const double* S) {
    return guarded([&] {
        require(ctx && S, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        ctx->check_replica(replica, false);
        const int64_t N = ctx->N();
        for (int64_t i = 0; i < 3 * N; ++i)
            require(std::isfinite(S[i]), PYRO_E_INVALID, "spin components must be finite");
        const size_t off = size_t(replica) * 3 * N;  // device spins S[r][i][a], the host layout
        if (ctx->prec == 64) {
            ctx->h2d(ctx->S.as<double>() + off, S, size_t(3 * N) * 8);
        } else {
            std::vector<float> v(3 * N);
            for (int64_t i = 0; i < 3 * N; ++i) v[i] = float(S[i]);
            ctx->h2d(ctx->S.as<float>() + off, v.data(), v.size() * 4);
        }
    });
}

int pyro_get_spins(pyro_ctx* ctx, int replica, double* S) {
    return guarded([&] {
        require(ctx && S, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        ctx->check_replica(replica, false);
        const int64_t N = ctx->N();
        const size_t off = size_t(replica) * 3 * N;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ctx->prec == 64) {
            ctx->d2h(S, ctx->S.as<double>() + off, size_t(3 * N) * 8);
        } else {
            std::vector<float> v(3 * N);
            ctx->d2h(v.data(), ctx->S.as<float>() + off, v.size() * 4);
            for (int64_t i = 0; i < 3 * N; ++i) S[i] = v[i];
        }
    });
}

int pyro_set_lifted_dir(pyro_ctx* ctx, int replica, const int8_t* dir) {
    return guarded([&] {
        require(ctx && dir, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        ctx->check_replica(replica, false);
        for (int64_t i = 0; i < ctx->N(); ++i)
            require(dir[i] == 1 || dir[i] == -1, PYRO_E_INVALID, "lifted directions must be +1 or -1");
        ctx->h2d(ctx->dir.as<int8_t>() + size_t(replica) * ctx->N(), dir, ctx->N());
    });
}

int pyro_get_lifted_dir(pyro_ctx* ctx, int replica, int8_t* dir) {
    return guarded([&] {
        require(ctx && dir, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        ctx->check_replica(replica, false);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->d2h(dir, ctx->dir.as<int8_t>() + size_t(replica) * ctx->N(), ctx->N());
    });
}

int pyro_randomize_spins(pyro_ctx* ctx, uint64_t seed) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        const uint32_t k0 = uint32_t(seed) ^ pyro::kTagInit, k1 = uint32_t(seed >> 32);
        if (ctx->prec == 64) HIPCHK(pyro::launch_randomize<double>(ctx->state<double>(nullptr), k0, k1, ctx->stream));
        else HIPCHK(pyro::launch_randomize<float>(ctx->state<float>(nullptr), k0, k1, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_pass_count(pyro_ctx* ctx, uint64_t* passes) {
    return guarded([&] {
        require(ctx && passes, PYRO_E_INVALID, "NULL argument");
        *passes = ctx->pass_count;
    });
}

int pyro_sweep(pyro_ctx* ctx, int kind, int n_overrelax, int n_sweeps, uint64_t* accepted) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        require(kind == PYRO_METROPOLIS || kind == PYRO_LIFTED || kind == PYRO_LIFTED_ROT, PYRO_E_INVALID,
                "sweep kind must be METROPOLIS, LIFTED or LIFTED_ROT");
        require(n_overrelax >= 0 && n_sweeps >= 0, PYRO_E_INVALID, "negative pass count");
        if (n_sweeps == 0) {
            if (accepted) std::fill(accepted, accepted + ctx->R, 0);
            return;
        }
        const int strat = ctx->choose_strategy(kind, n_overrelax);
        struct CountScope {  // counting on only for this call, also on error
            bool& f;
            CountScope(bool& flag, bool on) : f(flag) { f = on; }
            ~CountScope() { f = false; }
        } counting(ctx->count_acc, accepted != nullptr);
        if (accepted) {
            HIPCHK(pyro::launch_zero(ctx->acc.p, size_t(2) * ctx->R, ctx->stream));
            HIPCHK(pyro::launch_zero(ctx->acc_wave.p, size_t(ctx->R) * ctx->acc_waves, ctx->stream));
        }
        ctx->run_sweeps(strat, kind, n_overrelax, n_sweeps, nullptr, ctx->pass_count);
        ctx->pass_count += uint64_t(n_sweeps) * (1 + n_overrelax);
        if (accepted) {
            HIPCHK(pyro::launch_reduce_accepts(ctx->acc_wave.as<unsigned int>(), ctx->acc_waves,
                                               ctx->acc.as<unsigned long long>(), ctx->R, ctx->stream));
            std::vector<unsigned long long> a(ctx->R);
            HIPCHK(hipMemcpyAsync(a.data(), ctx->acc.p, sizeof(unsigned long long) * ctx->R,
                                  hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            for (int r = 0; r < ctx->R; ++r) accepted[r] = a[r];
        }
    });
}

int pyro_overrelax(pyro_ctx* ctx, int n_passes) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        require(n_passes >= 0, PYRO_E_INVALID, "negative pass count");
        const bool use_spec = ctx->strategy == 0 ? ctx->spec : ctx->strategy >= 3;
        for (int p = 0; p < n_passes; ++p) {
            if (ctx->prec == 64) ctx->pass_colours<double>(3, nullptr, ctx->pass_count, use_spec);
            else ctx->pass_colours<float>(3, nullptr, ctx->pass_count, use_spec);
            ctx->pass_count += 1;
        }
    });
}

int pyro_local_fields(pyro_ctx* ctx, int replica, double* h) {
    return guarded([&] {
        require(ctx && h, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        ctx->check_replica(replica, false);
        const int64_t N = ctx->N();
        const bool sten = ctx->stencil_observables();
        if (ctx->prec == 64) {
            if (sten)
                HIPCHK(pyro::launch_pyro_fields<double>(ctx->state<double>(nullptr), ctx->pyro_couplings<double>(),
                                                        replica, ctx->fields.as<double>(), ctx->stream));
            else
                HIPCHK(pyro::launch_fields<double>(ctx->state<double>(nullptr), replica, ctx->fields.as<double>(),
                                                   ctx->stream));
            std::vector<double> f(3 * N);
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->d2h(f.data(), ctx->fields.p, f.size() * 8);
            for (int64_t i = 0; i < N; ++i)
                for (int a = 0; a < 3; ++a) h[3 * i + a] = f[a * N + i];
        } else {
            if (sten)
                HIPCHK(pyro::launch_pyro_fields<float>(ctx->state<float>(nullptr), ctx->pyro_couplings<float>(),
                                                       replica, ctx->fields.as<float>(), ctx->stream));
            else
                HIPCHK(pyro::launch_fields<float>(ctx->state<float>(nullptr), replica, ctx->fields.as<float>(),
                                                  ctx->stream));
            std::vector<float> f(3 * N);
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->d2h(f.data(), ctx->fields.p, f.size() * 4);
            for (int64_t i = 0; i < N; ++i)
                for (int a = 0; a < 3; ++a) h[3 * i + a] = f[a * N + i];
        }
    });
}

int pyro_energy_per_cell(pyro_ctx* ctx, double* e) {
    return guarded([&] {
        require(ctx && e, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        ctx->energy(false);
        HIPCHK(hipMemcpyAsync(e, ctx->e.p, sizeof(double) * ctx->R, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_energy_per_cell_noalloc(const void* spins, int precision, const int32_t L[3], int n_replicas, double J1,
                                 double J2, double J3, double Delta, const double B[3], double* e, void* stream) {
    return guarded([&] {
        require(spins && L && B && e, PYRO_E_INVALID, "NULL argument");
        require(precision == PYRO_FP32 || precision == PYRO_FP64, PYRO_E_INVALID, "precision must be 32 or 64");
        require(n_replicas >= 1, PYRO_E_INVALID, "need at least one replica");
        for (double v : {J1, J2, J3, Delta, B[0], B[1], B[2]})
            require(std::isfinite(v), PYRO_E_INVALID, "couplings and field must be finite");
        require(J1 != 0.0, PYRO_E_UNSUPPORTED, "the stencil energy kernels need J1 != 0");
        int Ls[3];
        int64_t nc = 1;
        for (int a = 0; a < 3; ++a) {
            require(L[a] >= 1 && L[a] <= 1024, PYRO_E_INVALID, "lattice extent out of range");
            Ls[a] = L[a];
            nc *= L[a];
        }
        const int64_t N = 16 * nc;
        require(N <= (int64_t(1) << 30), PYRO_E_INVALID, "lattice too large");
        int device = 0;
        HIPCHK(hipGetDevice(&device));
        double* scr = noalloc_scratch(device);
        const size_t rs = precision == PYRO_FP64 ? 8 : 4;
        // the kernel pyro_energy_per_cell would use for this lattice
        const int ek = energy_tile_kind_for(Ls, rs, g_scratch_lds[device], auto_tile_kind(Ls, rs, g_scratch_lds[device]));
        int64_t nblk = (N + 63) / 64;
        if (ek) {
            int d[3];
            pyro::pyro_tile_dims(ek, d);
            nblk = int64_t(Ls[0] / d[0]) * (Ls[1] / d[1]) * (Ls[2] / d[2]);
        }
        // partials [R][nblk], then the field [R][3] on a 256-byte boundary
        const int64_t R = n_replicas;
        const int64_t b_off = (R * nblk + 31) / 32 * 32;
        require(b_off + 3 * R <= int64_t(kNoAllocScratchDoubles), PYRO_E_INVALID,
                "replicas x energy blocks exceed the allocation-free scratch (4 MiB)");
        auto run = [&](auto real) {
            using Real = decltype(real);
            Real* Bd = reinterpret_cast<Real*>(scr + b_off);
            pyro::DeviceState<Real> st;
            st.S = static_cast<Real*>(const_cast<void*>(spins));  // read only
            st.B = Bd;
            st.N = int(N);
            st.n_cells = int(nc);
            st.R = n_replicas;
            st.L0 = Ls[0];
            st.L1 = Ls[1];
            st.L2 = Ls[2];
            pyro::PyroCouplings<Real> J;
            J.J1 = Real(J1);
            J.J2 = Real(J2);
            J.J3 = Real(J3);
            J.Jx = Real(J1 * (Delta - 1.0) / 3.0);
            J.shells = J3 != 0.0 ? 3 : (J2 != 0.0 ? 2 : 1);
            J.xxz = Delta != 1.0;
            const hipStream_t s = static_cast<hipStream_t>(stream);
            HIPCHK(pyro::launch_fill_field<Real>(Bd, n_replicas, B[0], B[1], B[2], s));
            int nb = int(nblk);
            if (ek) HIPCHK(pyro::launch_pyro_energy_tiled<Real>(st, J, ek, scr, &nb, s));
            else HIPCHK(pyro::launch_pyro_energy<Real>(st, J, scr, s));
            HIPCHK(pyro::launch_energy_finish(scr, nb, int(nc), e, nullptr, nullptr, nullptr, nullptr, n_replicas, s));
        };
        if (precision == PYRO_FP64) run(double{});
        else run(float{});
    });
}

int pyro_energy_sample(pyro_ctx* ctx) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        ctx->energy(true);
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_pt_init(pyro_ctx* ctx, const double* T_ladder, uint64_t swap_seed) {
    return guarded([&] {
        require(ctx && T_ladder, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        for (int r = 0; r < ctx->R; ++r)
            require(T_ladder[r] > 0 && std::isfinite(T_ladder[r]), PYRO_E_INVALID, "temperatures must be positive");
        ctx->h_T.assign(T_ladder, T_ladder + ctx->R);
        ctx->upload_temperatures();
        ctx->identity_permutation();
        ctx->swap_k0 = uint32_t(swap_seed) ^ pyro::kTagSwap;
        ctx->swap_k1 = uint32_t(swap_seed >> 32);
        ctx->round_count = 0;
        HIPCHK(pyro::launch_zero(ctx->Esum.p, size_t(2) * ctx->R, ctx->stream));
        HIPCHK(pyro::launch_zero(ctx->E2sum.p, size_t(2) * ctx->R, ctx->stream));
        HIPCHK(pyro::launch_zero(ctx->nE.p, size_t(2) * ctx->R, ctx->stream));
        HIPCHK(pyro::launch_zero(ctx->att.p, size_t(2) * ctx->R, ctx->stream));
        HIPCHK(pyro::launch_zero(ctx->accp.p, size_t(2) * ctx->R, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->pt_ready = true;
        ctx->invalidate_graph();
    });
}

int pyro_pt_rounds(pyro_ctx* ctx, int kind, int n_overrelax, int n_sweep, int sample_ssf,
                   int n_rounds) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        require(ctx->pt_ready, PYRO_E_STATE, "pt_init() has not been called");
        require(kind == PYRO_METROPOLIS || kind == PYRO_LIFTED || kind == PYRO_LIFTED_ROT, PYRO_E_INVALID,
                "sweep kind must be METROPOLIS, LIFTED or LIFTED_ROT");
        require(n_overrelax >= 0 && n_sweep >= 0 && n_rounds >= 0, PYRO_E_INVALID, "negative count");
        require(sample_ssf >= 0 && sample_ssf <= 2, PYRO_E_INVALID, "sample_ssf must be 0, 1 or 2");
        require(!sample_ssf || ctx->ssf_ready, PYRO_E_STATE, "ssf_configure() has not been called");
        require(!sample_ssf || ctx->has_ssf_map || ctx->n_slots >= ctx->R, PYRO_E_INVALID,
                "SSF sampling in tempering rounds needs a slot map or one accumulator slot per replica");
        // SSF samples per round: one after the sweeps, or one after each sweep
        const int spr = sample_ssf == 2 ? n_sweep : (sample_ssf ? 1 : 0);
        require(!ctx->munu || spr <= ctx->munu_B, PYRO_E_INVALID,
                "a round's SSF samples exceed the sublattice-resolved buffer");
        if (n_rounds == 0) return;
        const int strat = n_sweep > 0 ? ctx->choose_strategy(kind, n_overrelax) : 1;
        const unsigned long long ppr = (unsigned long long)n_sweep * (1 + n_overrelax);
        auto* ctr = ctx->counters.as<unsigned long long>();
        // device counters: pass base, round
        unsigned long long hc[2] = {ctx->pass_count, ctx->round_count};
        HIPCHK(hipMemcpyAsync(ctx->counters.p, hc, sizeof(hc), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));

        // the round's last colour pass leaves the energy partials when it can
        const bool fused_energy = n_sweep > 0 && ctx->sweeps_leave_energy(strat);
        auto ssf_round_sample = [&]() {
            if (ctx->has_ssf_map) {
                HIPCHK(pyro::launch_map_slots(ctx->slot_of.as<int32_t>(), ctx->ssf_map.as<int32_t>(),
                                              ctx->ssf_slot.as<int32_t>(), ctx->R, ctx->stream));
                ctx->ssf_sample_device(ctx->ssf_slot.as<int32_t>(), ctx->stream);
            } else {
                ctx->ssf_sample_device(ctx->slot_of.as<int32_t>(), ctx->stream);
            }
        };
        // the round's energy sample and exchange, after its sweeps and SSF
        auto energy_exchange = [&]() {
            if (fused_energy) ctx->energy_finish(true, pyro::pyro_colour_energy_blocks(int(ctx->nc())));
            else ctx->energy(true);
            pyro::ExchangeArgs a;
            a.e = ctx->e.as<double>();
            a.rep_of = ctx->rep_of.as<int32_t>();
            a.slot_of = ctx->slot_of.as<int32_t>();
            a.T_slot = ctx->T_slot.as<double>();
            a.att = ctx->att.as<unsigned long long>();
            a.accp = ctx->accp.as<unsigned long long>();
            a.round = ctr + 1;
            a.pass_base = ctr;
            a.passes_per_round = ppr;
            a.every = (unsigned long long)ctx->exchange_every;
            a.n_cells = double(ctx->nc());
            a.k0 = ctx->swap_k0;
            a.k1 = ctx->swap_k1;
            a.R = ctx->R;
            ctx->timed(kFamExchange, [&] { return pyro::launch_exchange(a, ctx->stream); });
        };
        auto round_body = [&]() {
            if (sample_ssf == 2) {
                // an SSF sample after every sweep of the round
                for (int sw = 0; sw < n_sweep; ++sw) {
                    ctx->run_sweeps(strat, kind, n_overrelax, 1, ctr, (unsigned long long)sw * (1 + n_overrelax),
                                    fused_energy && sw == n_sweep - 1 ? ctx->epart.as<double>() : nullptr);
                    ssf_round_sample();
                }
            } else {
                if (n_sweep > 0)
                    ctx->run_sweeps(strat, kind, n_overrelax, n_sweep, ctr, 0ull,
                                    fused_energy ? ctx->epart.as<double>() : nullptr);
                // One stream.  The SSF on a second stream beside the energy
                // (fork and join by events inside the captured round) made each
                // round 0.19 ms slower; sampled from a snapshot beside the next
                // round's XCD-resident sweep it slowed that sweep by more than it
                // saved, and the snapshot copy cost as much again.
                if (sample_ssf) ssf_round_sample();
            }
            energy_exchange();
        };
        const bool use_graph = !ctx->prof;
        if (use_graph) {
            if (!ctx->gexec || ctx->g_kind != kind || ctx->g_nor != n_overrelax || ctx->g_nsw != n_sweep ||
                ctx->g_ssf != sample_ssf || ctx->g_strategy != strat) {
                ctx->invalidate_graph();
                ctx->gexec = capture_graph(ctx->stream, round_body);
                ctx->g_kind = kind;
                ctx->g_nor = n_overrelax;
                ctx->g_nsw = n_sweep;
                ctx->g_ssf = sample_ssf;
                ctx->g_strategy = strat;
            }
        }
        for (int i = 0; i < n_rounds; ++i) {
            // flush the buffered rows before a round that would not fit
            if (spr && ctx->munu && ctx->munu_pending + spr > ctx->munu_B) ctx->munu_flush();
            if (use_graph) HIPCHK(hipGraphLaunch(ctx->gexec, ctx->stream));
            else round_body();
            ctx->pass_count += ppr;
            ctx->round_count += 1;
            if (spr && ctx->munu) {
                ctx->munu_pending += spr;
                if (ctx->munu_pending >= ctx->munu_B) ctx->munu_flush();
            }
        }
    });
}

int pyro_set_replica_couplings(pyro_ctx* ctx, const double* J1J2J3Delta) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        require(ctx->spec, PYRO_E_UNSUPPORTED, "per-replica couplings need couplings defined by define_J1J2J3");
        ctx->invalidate_graph();
        if (!J1J2J3Delta) {
            ctx->has_Jrep = false;
            ctx->xxz_rep = false;
            ctx->Jrep.release();
            return;
        }
        require(ctx->strategy == 0 || ctx->strategy >= 3, PYRO_E_UNSUPPORTED,
                "per-replica couplings need the pyrochlore strategies 3, 4 or 5");
        const int R = ctx->R;
        std::vector<double> v(size_t(4) * R);
        bool xxz = false;
        for (int r = 0; r < R; ++r) {
            const double* q = J1J2J3Delta + 4 * r;
            for (int k = 0; k < 4; ++k) require(std::isfinite(q[k]), PYRO_E_INVALID, "couplings must be finite");
            // the colouring was built for the context's shells: a replica may
            // not add a shell the context left out
            require(ctx->spec_shells >= 2 || q[1] == 0.0, PYRO_E_INVALID,
                    "replica J2 != 0 but the context was defined with J2 = 0");
            require(ctx->spec_shells >= 3 || q[2] == 0.0, PYRO_E_INVALID,
                    "replica J3 != 0 but the context was defined with J3 = 0");
            v[4 * r] = q[0];
            v[4 * r + 1] = q[1];
            v[4 * r + 2] = q[2];
            v[4 * r + 3] = q[0] * (q[3] - 1.0) / 3.0;
            xxz = xxz || q[3] != 1.0;
        }
        const size_t rs = ctx->real_size();
        ctx->Jrep.alloc(rs * 4 * R);
        if (rs == 8) {
            ctx->h2d(ctx->Jrep.p, v.data(), 8 * v.size());
        } else {
            std::vector<float> f(v.begin(), v.end());
            ctx->h2d(ctx->Jrep.p, f.data(), 4 * f.size());
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->has_Jrep = true;
        ctx->xxz_rep = xxz;
    });
}

int pyro_anneal_schedule(pyro_ctx* ctx, const double* T, int n_steps) {
    return guarded([&] {
        require(ctx && T, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        require(n_steps > 0, PYRO_E_INVALID, "the schedule needs at least one step");
        for (int k = 0; k < n_steps; ++k)
            require(T[k] > 0 && std::isfinite(T[k]), PYRO_E_INVALID, "temperatures must be positive");
        ctx->invalidate_graph();
        ctx->sched.alloc(sizeof(double) * n_steps);
        ctx->ehist.alloc(sizeof(double) * size_t(n_steps) * ctx->R);
        ctx->h2d(ctx->sched.p, T, sizeof(double) * n_steps);
        HIPCHK(pyro::launch_zero(ctx->ehist.p, size_t(2) * n_steps * ctx->R, ctx->stream));
        ctx->n_sched = n_steps;
        ctx->h_sched.assign(T, T + n_steps);
        ctx->anneal_step = 0;
        // every replica is its own model: identity permutation, slot r = replica r
        ctx->h_T.assign(ctx->R, T[0]);
        ctx->upload_temperatures();
        ctx->identity_permutation();
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_anneal_steps(pyro_ctx* ctx, int kind, int n_overrelax, int n_sweep, int sample_ssf, int n_steps) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->anneal_check(kind, n_overrelax, n_sweep, sample_ssf, n_steps);
        if (n_steps == 0) return;
        const int strat = n_sweep > 0 ? ctx->choose_strategy(kind, n_overrelax) : 1;
        const unsigned long long ppr = (unsigned long long)n_sweep * (1 + n_overrelax);
        ctx->anneal_upload_counters();
        // the step's last colour pass leaves the energy partials when it can
        const bool fused_energy = n_sweep > 0 && ctx->sweeps_leave_energy(strat);
        auto step_body = [&]() {
            ctx->anneal_set_T();
            if (n_sweep > 0)
                ctx->run_sweeps(strat, kind, n_overrelax, n_sweep, ctx->counters.as<unsigned long long>(), 0ull,
                                fused_energy ? ctx->epart.as<double>() : nullptr);
            ctx->anneal_observe(sample_ssf, fused_energy, ppr);
        };
        const bool use_graph = !ctx->prof;
        if (use_graph && (!ctx->agexec || ctx->a_kind != kind || ctx->a_nor != n_overrelax ||
                          ctx->a_nsw != n_sweep || ctx->a_ssf != sample_ssf || ctx->a_strategy != strat)) {
            if (ctx->agexec) (void)hipGraphExecDestroy(ctx->agexec);
            ctx->agexec = nullptr;
            ctx->agexec = capture_graph(ctx->stream, step_body);
            ctx->a_kind = kind;
            ctx->a_nor = n_overrelax;
            ctx->a_nsw = n_sweep;
            ctx->a_ssf = sample_ssf;
            ctx->a_strategy = strat;
        }
        // enqueue only: contexts on their own streams run concurrently
        for (int i = 0; i < n_steps; ++i) {
            if (use_graph) HIPCHK(hipGraphLaunch(ctx->agexec, ctx->stream));
            else step_body();
            if (ctx->anneal_advance(sample_ssf, ppr)) ctx->munu_flush();
        }
        ctx->h_T.assign(ctx->R, ctx->h_sched[ctx->anneal_step - 1]);
    });
}

int pyro_anneal_history(pyro_ctx* ctx, double* E) {
    return guarded([&] {
        require(ctx && E, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        require(ctx->n_sched > 0, PYRO_E_STATE, "anneal_schedule() has not been called");
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->d2h(E, ctx->ehist.p, sizeof(double) * size_t(ctx->n_sched) * ctx->R);
    });
}

int pyro_pt_stats(pyro_ctx* ctx, uint64_t* attempts, uint64_t* accepts, double* E, double* E2,
                  uint64_t* n) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const int R = ctx->R;
        auto get_u64 = [&](const DevMem& m, uint64_t* dst, int cnt) {
            if (!dst || cnt <= 0) return;
            std::vector<unsigned long long> v(cnt);
            ctx->d2h(v.data(), m.p, sizeof(unsigned long long) * cnt);
            for (int i = 0; i < cnt; ++i) dst[i] = v[i];
        };
        get_u64(ctx->att, attempts, R - 1);
        get_u64(ctx->accp, accepts, R - 1);
        get_u64(ctx->nE, n, R);
        if (E) ctx->d2h(E, ctx->Esum.p, sizeof(double) * R);
        if (E2) ctx->d2h(E2, ctx->E2sum.p, sizeof(double) * R);
    });
}

int pyro_pt_permutation(pyro_ctx* ctx, int32_t* replica_of_slot) {
    return guarded([&] {
        require(ctx && replica_of_slot, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->d2h(replica_of_slot, ctx->rep_of.p, sizeof(int32_t) * ctx->R);
    });
}

int pyro_pt_reset_stats(pyro_ctx* ctx) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        for (DevMem* m : {&ctx->Esum, &ctx->E2sum, &ctx->nE, &ctx->att, &ctx->accp})
            HIPCHK(pyro::launch_zero(m->p, m->bytes / 4, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_pt_set_exchange_interval(pyro_ctx* ctx, int every) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(every >= 1, PYRO_E_INVALID, "the exchange interval must be at least 1");
        ctx->exchange_every = every;
        ctx->invalidate_graph();
    });
}

int pyro_ssf_configure(pyro_ctx* ctx, int n_corr, const int32_t* alpha, const int32_t* beta,
                       int n_slots, int store_error_term, int munu, int munu_buffer) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        require(n_corr > 0 && alpha && beta, PYRO_E_INVALID, "need at least one correlator");
        require(n_slots > 0, PYRO_E_INVALID, "need at least one accumulator slot");
        for (int c = 0; c < n_corr; ++c)
            require(alpha[c] >= 0 && alpha[c] < 3 && beta[c] >= 0 && beta[c] < 3, PYRO_E_INVALID,
                    "correlator components must be 0 (x), 1 (y) or 2 (z)");
        const int n_f = 3 * ctx->geom.n_sl;
        if (munu) {
            require(munu_buffer > 0, PYRO_E_INVALID, "munu_buffer must be positive");
            require(n_f <= 48, PYRO_E_UNSUPPORTED, "sublattice-resolved products support at most 16 sublattices");
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->invalidate_graph();
        ctx->n_corr = n_corr;
        ctx->h_alpha.assign(alpha, alpha + n_corr);
        ctx->h_beta.assign(beta, beta + n_corr);
        ctx->n_slots = n_slots;
        ctx->store2 = store_error_term != 0;
        ctx->munu = munu != 0;
        ctx->munu_B = munu ? munu_buffer : 0;
        const size_t rs = ctx->real_size();
        const int64_t nc = ctx->nc();
        const int n_sl = ctx->geom.n_sl;
        ctx->ensure_dft();
        ctx->d_alpha.alloc(sizeof(int32_t) * n_corr);
        ctx->d_beta.alloc(sizeof(int32_t) * n_corr);
        ctx->h2d(ctx->d_alpha.p, alpha, sizeof(int32_t) * n_corr);
        ctx->h2d(ctx->d_beta.p, beta, sizeof(int32_t) * n_corr);
        const size_t acc_bytes = sizeof(double) * 2 * size_t(n_slots) * n_corr * nc;
        ctx->sacc.alloc(acc_bytes);
        HIPCHK(hipMemset(ctx->sacc.p, 0, acc_bytes));
        if (ctx->store2) {
            ctx->sacc2.alloc(acc_bytes);
            HIPCHK(hipMemset(ctx->sacc2.p, 0, acc_bytes));
        } else {
            ctx->sacc2.release();
        }
        ctx->sn.alloc(sizeof(unsigned long long) * n_slots);
        HIPCHK(hipMemset(ctx->sn.p, 0, ctx->sn.bytes));
        ctx->ssf_slot.alloc(sizeof(int32_t) * ctx->R);
        if (ctx->munu) {
            ctx->munu_T = (n_f + 15) / 16;
            const size_t fw = size_t(32) * ctx->munu_T;
            ctx->mbuf.alloc(sizeof(float) * fw * size_t(n_slots) * ctx->n_kh * ctx->munu_B);
            ctx->mcnt.alloc(sizeof(unsigned int) * n_slots);
            HIPCHK(hipMemset(ctx->mcnt.p, 0, ctx->mcnt.bytes));
            // the stored tile pairs only (Hermitian: 6 of the 9 tiles at 48 features)
            // (and the kept K only: M(-K) = conj(M(K)))
            const size_t mbytes =
                sizeof(double) * 2 * size_t(n_slots) * ctx->n_kh * pyro::munu_acc_pairs(ctx->munu_T) * 256;
            ctx->macc.alloc(mbytes);
            HIPCHK(hipMemset(ctx->macc.p, 0, mbytes));
            ctx->host_munu_count.assign(n_slots, 0);
            // FP32: the contraction buffers its products beside the rows and
            // the flush adds them to the FP64 sums (no per-sample update)
            if (ctx->prec == 32)
                ctx->cbuf.alloc(sizeof(float) * 2 * size_t(n_slots) * ctx->munu_B * n_corr * ctx->n_kh);
            else
                ctx->cbuf.release();
        } else {
            ctx->mbuf.release();
            ctx->mcnt.release();
            ctx->macc.release();
            ctx->cbuf.release();
        }
        HIPCHK(hipDeviceSynchronize());
        ctx->munu_pending = 0;
        ctx->has_ssf_map = false;  // a map names slots of the previous configuration
        ctx->ssf_ready = true;
    });
}

int pyro_pt_ssf_map(pyro_ctx* ctx, const int32_t* map) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(ctx->ssf_ready, PYRO_E_STATE, "ssf_configure() has not been called");
        if (!map) {
            if (ctx->has_ssf_map) ctx->invalidate_graph();
            ctx->has_ssf_map = false;
            return;
        }
        std::vector<int> used(ctx->n_slots, 0);
        for (int s = 0; s < ctx->R; ++s) {
            require(map[s] >= -1 && map[s] < ctx->n_slots, PYRO_E_INVALID, "slot map entry out of range");
            // one rung per accumulator: the accumulation kernels of one
            // round update each slot from a single replica
            if (map[s] >= 0) require(used[map[s]]++ == 0, PYRO_E_INVALID, "two rungs map to one SSF slot");
        }
        // the map lives at a fixed device address that captured rounds and
        // anneal steps read when replayed, so a new map needs no new graph
        if (!ctx->ssf_map.p) ctx->ssf_map.alloc(sizeof(int32_t) * ctx->R);
        ctx->h2d(ctx->ssf_map.p, map, sizeof(int32_t) * ctx->R);
        if (!ctx->has_ssf_map) ctx->invalidate_graph();
        ctx->has_ssf_map = true;
    });
}

int pyro_ssf_sample(pyro_ctx* ctx, const int32_t* slot_of_replica) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        require(ctx->ssf_ready, PYRO_E_STATE, "ssf_configure() has not been called");
        if (!slot_of_replica) {
            require(ctx->n_slots >= ctx->R, PYRO_E_INVALID, "tempering slots exceed the accumulator slots");
            ctx->ssf_sample_device(ctx->slot_of.as<int32_t>(), ctx->stream);
            if (ctx->munu) {
                for (auto& c : ctx->host_munu_count) ++c;
                if (++ctx->munu_pending >= ctx->munu_B) ctx->munu_flush();
            }
            HIPCHK(hipStreamSynchronize(ctx->stream));
            return;
        }
        for (int r = 0; r < ctx->R; ++r)
            require(slot_of_replica[r] < ctx->n_slots, PYRO_E_INVALID, "slot index out of range");
        // Split into layers in which every slot appears at most once, so the
        // accumulators see no concurrent updates.
        std::vector<int> seen(ctx->n_slots, 0);
        int n_layers = 0;
        std::vector<int> layer(ctx->R, -1);
        for (int r = 0; r < ctx->R; ++r) {
            const int s = slot_of_replica[r];
            if (s < 0) continue;
            layer[r] = seen[s]++;
            n_layers = std::max(n_layers, layer[r] + 1);
        }
        for (int l = 0; l < n_layers; ++l) {
            std::vector<int32_t> sl(ctx->R, -1);
            for (int r = 0; r < ctx->R; ++r)
                if (layer[r] == l) sl[r] = slot_of_replica[r];
            if (ctx->munu) {
                // flush slots whose buffer is full before staging more
                bool full = false;
                for (int r = 0; r < ctx->R; ++r)
                    if (sl[r] >= 0 && ctx->host_munu_count[sl[r]] >= ctx->munu_B) full = true;
                if (full) ctx->munu_flush();
            }
            HIPCHK(hipMemcpyAsync(ctx->ssf_slot.p, sl.data(), sizeof(int32_t) * ctx->R, hipMemcpyHostToDevice,
                                  ctx->stream));
            ctx->ssf_sample_device(ctx->ssf_slot.as<int32_t>(), ctx->stream);
            if (ctx->munu)
                for (int r = 0; r < ctx->R; ++r)
                    if (sl[r] >= 0) ctx->host_munu_count[sl[r]] += 1;
            HIPCHK(hipStreamSynchronize(ctx->stream));
        }
    });
}

int pyro_ssf_read(pyro_ctx* ctx, int slot, double* S, double* S2, uint64_t* n_samples) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(ctx->ssf_ready, PYRO_E_STATE, "ssf_configure() has not been called");
        require(slot >= 0 && slot < ctx->n_slots, PYRO_E_INVALID, "slot index out of range");
        require(!S2 || ctx->store2, PYRO_E_STATE, "error term was not configured");
        ctx->munu_flush();  // buffered contracted samples into the sums
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const int64_t nc = ctx->nc();
        const size_t per = size_t(2) * ctx->n_corr * nc;
        if (S) ctx->d2h(S, ctx->sacc.as<double>() + slot * per, per * 8);
        if (S2) ctx->d2h(S2, ctx->sacc2.as<double>() + slot * per, per * 8);
        // K outside the kept half: S(K) = conj(S(-K)) per sample (b(K) =
        // conj(b(-K)) wherever the centred q of K is minus that of -K); the
        // squares of the parts are equal
        for (int64_t K = 0; K < nc; ++K) {
            if (ctx->kept_index(K) >= 0 || ctx->nyquist_01(K)) continue;
            const int64_t P = ctx->neg_k(K);
            for (int c = 0; c < ctx->n_corr; ++c) {
                const size_t d = 2 * (size_t(c) * nc + K), q = 2 * (size_t(c) * nc + P);
                if (S) {
                    S[d] = S[q];
                    S[d + 1] = -S[q + 1];
                }
                if (S2) {
                    S2[d] = S2[q];
                    S2[d + 1] = S2[q + 1];
                }
            }
        }
        if (n_samples) {
            unsigned long long n = 0;
            ctx->d2h(&n, ctx->sn.as<unsigned long long>() + slot, 8);
            *n_samples = n;
        }
    });
}

int pyro_ssf_read_munu(pyro_ctx* ctx, int slot, int c, double* C) {
    return guarded([&] {
        require(ctx && C, PYRO_E_INVALID, "NULL argument");
        require(ctx->ssf_ready && ctx->munu, PYRO_E_STATE, "sublattice-resolved products were not configured");
        require(slot >= 0 && slot < ctx->n_slots, PYRO_E_INVALID, "slot index out of range");
        require(c >= 0 && c < ctx->n_corr, PYRO_E_INVALID, "correlator index out of range");
        ctx->munu_flush();
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const int n_sl = ctx->geom.n_sl, T = ctx->munu_T;
        const int64_t nc = ctx->nc();
        const size_t per_k = size_t(pyro::munu_acc_pairs(T)) * 256 * 2;  // doubles
        std::vector<double> packed(size_t(ctx->n_kh) * per_k);
        ctx->d2h(packed.data(), ctx->macc.as<double>() + size_t(slot) * packed.size(), packed.size() * 8);
        const int al = ctx->h_alpha[c], be = ctx->h_beta[c];
        // The flush accumulates 16-feature tile pairs (I, J) with I <= J
        // only; an entry of a lower tile is the conjugate of its transpose
        // (the matrix sum_samples conj(A_f) A_g is Hermitian).  K outside the
        // kept half: M(K) = conj(M(-K)), since A(K) = conj(A(-K)).
        for (int64_t K = 0; K < nc; ++K) {
            int64_t kh = ctx->kept_index(K);
            const bool outside = kh < 0;
            if (outside) kh = ctx->kept_index(ctx->neg_k(K));
            for (int mu = 0; mu < n_sl; ++mu)
                for (int nu = 0; nu < n_sl; ++nu) {
                    int f = al * n_sl + mu, g = be * n_sl + nu;
                    const bool lower = f / 16 > g / 16;
                    if (lower) std::swap(f, g);
                    const size_t src = size_t(kh) * per_k +
                                       (size_t(pyro::munu_acc_pair(f / 16, g / 16, T)) * 256 + (f % 16) * 16 + g % 16) * 2;
                    const size_t dst = ((size_t(K) * n_sl + mu) * n_sl + nu) * 2;
                    C[dst] = packed[src];
                    C[dst + 1] = lower != outside ? -packed[src + 1] : packed[src + 1];
                }
        }
    });
}

int pyro_ssf_reset(pyro_ctx* ctx, int slot) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(ctx->ssf_ready, PYRO_E_STATE, "ssf_configure() has not been called");
        require(slot >= -1 && slot < ctx->n_slots, PYRO_E_INVALID, "slot index out of range");
        const size_t per = size_t(2) * ctx->n_corr * ctx->nc();
        const int s0 = slot < 0 ? 0 : slot, s1 = slot < 0 ? ctx->n_slots : slot + 1;
        ctx->munu_flush();
        for (int s = s0; s < s1; ++s) {
            HIPCHK(pyro::launch_zero(ctx->sacc.as<double>() + s * per, per * 2, ctx->stream));
            if (ctx->store2) HIPCHK(pyro::launch_zero(ctx->sacc2.as<double>() + s * per, per * 2, ctx->stream));
            HIPCHK(pyro::launch_zero(ctx->sn.as<unsigned long long>() + s, 2, ctx->stream));
            if (ctx->munu) {
                const size_t mper = size_t(2) * ctx->n_kh * pyro::munu_acc_pairs(ctx->munu_T) * 256;
                HIPCHK(pyro::launch_zero(ctx->macc.as<double>() + s * mper, mper * 2, ctx->stream));
            }
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_spin_ft(pyro_ctx* ctx, int replica, double* ft) {
    return guarded([&] {
        require(ctx && ft, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        ctx->check_replica(replica, false);
        ctx->ensure_dft();
        const int64_t nc = ctx->nc(), nh = ctx->n_kh;
        const size_t n = size_t(3) * ctx->geom.n_sl * nh * 2;  // the kept half
        std::vector<double> h(n);
        if (ctx->prec == 64) {
            HIPCHK(pyro::launch_dft<double>(ctx->ssf_state<double>(), replica, 1, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->d2h(h.data(), ctx->F.p, n * 8);
        } else {
            HIPCHK(pyro::launch_dft<float>(ctx->ssf_state<float>(), replica, 1, ctx->stream));
            std::vector<float> f(n);
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->d2h(f.data(), ctx->F.p, n * 4);
            for (size_t i = 0; i < n; ++i) h[i] = f[i];
        }
        // every K: kept, or the conjugate of -K (real fields)
        for (int64_t K = 0; K < nc; ++K) {
            int64_t kh = ctx->kept_index(K);
            const double sg = kh < 0 ? -1.0 : 1.0;
            if (kh < 0) kh = ctx->kept_index(ctx->neg_k(K));
            for (int f = 0; f < 3 * ctx->geom.n_sl; ++f) {
                ft[2 * (size_t(f) * nc + K)] = h[2 * (size_t(f) * nh + kh)];
                ft[2 * (size_t(f) * nc + K) + 1] = sg * h[2 * (size_t(f) * nh + kh) + 1];
            }
        }
    });
}

int pyro_sweep_order(pyro_ctx* ctx, int n_overrelax, int32_t* site, int32_t* pass_offset) {
    return guarded([&] {
        require(ctx && site && pass_offset, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        require(n_overrelax >= 0, PYRO_E_INVALID, "negative pass count");
        const int strat = ctx->strategy != 0 ? ctx->strategy
                          : (ctx->tile_kind > 0 && ctx->nc() >= 4096) ? 5
                          : (ctx->spec ? 3 : 1);
        const auto& L = ctx->geom.L;
        const int64_t nc = ctx->nc();
        const auto& t = ctx->topo;
        const int P = t.period[0] * t.period[1] * t.period[2];
        int64_t o = 0;
        auto class_of = [&](int64_t i) {
            const int s = int(i / nc);
            int64_t c = i % nc;
            const int n2 = int(c % L[2]);
            c /= L[2];
            const int n1 = int(c % L[1]);
            const int n0 = int(c / L[1]);
            return s * P + ((n0 % t.period[0]) * t.period[1] + (n1 % t.period[1])) * t.period[2] + (n2 % t.period[2]);
        };
        // the fused pass order: stage group 0 gives each site its first
        // kind and its first OR update back to back, group g >= 1 OR pass
        // g + 1 alone; otherwise group k is pass k alone
        const bool fused = ctx->fused_order(strat, n_overrelax);
        const int n_groups = fused ? std::max(n_overrelax, 1) : n_overrelax + 1;
        auto emit = [&](int grp, int32_t i) {
            site[o] = i;
            pass_offset[o++] = fused && grp > 0 ? grp + 1 : grp;
            if (fused && grp == 0) {
                site[o] = i;
                pass_offset[o++] = 1;
            }
        };
        if (strat != 5) {
            for (int g = 0; g < n_groups; ++g)
                for (int c = 0; c < t.n_colours; ++c)
                    for (int64_t i = 0; i < ctx->N(); ++i)
                        if (t.class_colour[class_of(i)] == c) emit(g, int32_t(i));
            return;
        }
        require(ctx->tile_kind > 0, PYRO_E_UNSUPPORTED, "tiled strategy does not apply");
        int T[3];
        pyro::pyro_tile_dims(ctx->tile_kind, T);
        const int ncol = ctx->spec_shells == 3 ? 16 : 4;
        for (int phase = 0; phase < 8; ++phase)
            for (int m0 = 0; m0 < L[0] / (2 * T[0]); ++m0)
                for (int m1 = 0; m1 < L[1] / (2 * T[1]); ++m1)
                    for (int m2 = 0; m2 < L[2] / (2 * T[2]); ++m2) {
                        const int o0 = ((phase & 1) + 2 * m0) * T[0];
                        const int o1 = (((phase >> 1) & 1) + 2 * m1) * T[1];
                        const int o2 = (((phase >> 2) & 1) + 2 * m2) * T[2];
                        // the stage groups within the tile, as above
                        for (int g = 0; g < n_groups; ++g)
                            for (int col = 0; col < ncol; ++col)
                                for (int q = 0; q < 16 / ncol; ++q) {
                                    const int sl = ncol == 16 ? col : col + 4 * q;
                                    for (int i0 = 0; i0 < T[0]; ++i0)
                                        for (int i1 = 0; i1 < T[1]; ++i1)
                                            for (int i2 = 0; i2 < T[2]; ++i2)
                                                emit(g, int32_t(sl * nc + ctx->geom.cell_index(o0 + i0, o1 + i1, o2 + i2)));
                                }
                    }
    });
}

int pyro_synchronize(pyro_ctx* ctx) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int pyro_set_strategy(pyro_ctx* ctx, int strategy) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(strategy >= 0 && strategy <= 6, PYRO_E_INVALID, "strategy must be 0..6");
        ctx->strategy = strategy;
        ctx->tuned_key = -1;
        ctx->invalidate_graph();
    });
}

int pyro_set_wait_timeout(pyro_ctx* ctx, double seconds) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        require(seconds > 0.0 && seconds <= 60.0, PYRO_E_INVALID, "the wait bound must lie in (0, 60] seconds");
        int dev = 0, khz = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        const double ticks = seconds * 1e3 * double(khz > 0 ? khz : 100000);
        ctx->xcd_ticks = ticks < 1.0 ? 1ull : (unsigned long long)ticks;
        ctx->invalidate_graph();  // the bound is a kernel argument of captured rounds
    });
}

int pyro_set_fused_passes(pyro_ctx* ctx, int on) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        if (ctx->fused_passes != (on != 0)) {
            ctx->invalidate_graph();
            ctx->tuned_key = -1;  // time the candidates again in the new order
        }
        ctx->fused_passes = on != 0;
    });
}

int pyro_set_tile_shape(pyro_ctx* ctx, const int32_t T[3]) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->need_setup();
        int kind = 0;
        if (T) {
            for (int k = 1; k <= pyro::kPyroTileKinds && !kind; ++k) {
                int d[3];
                pyro::pyro_tile_dims(k, d);
                if (d[0] == T[0] && d[1] == T[1] && d[2] == T[2]) kind = k;
            }
            require(kind > 0, PYRO_E_UNSUPPORTED, "tile shape not available (4x4x4, 6x6x4, 8x4x4, 4x4x8, 8x8x4)");
            require(ctx->tile_applies(kind), PYRO_E_UNSUPPORTED,
                    "tile shape does not divide the lattice (L_a divisible by twice the edge) or exceeds LDS");
        }
        ctx->tile_forced = kind;
        ctx->tile_kind = ctx->pick_tile_kind();
        ctx->tuned_key = -1;
        ctx->invalidate_graph();
    });
}

int pyro_get_tile_shape(pyro_ctx* ctx, int32_t T[3]) {
    return guarded([&] {
        require(ctx && T, PYRO_E_INVALID, "NULL argument");
        ctx->need_setup();
        int d[3];
        pyro::pyro_tile_dims(ctx->tile_kind, d);
        for (int a = 0; a < 3; ++a) T[a] = d[a];
    });
}

int pyro_get_strategy(pyro_ctx* ctx, int* strategy) {
    return guarded([&] {
        require(ctx && strategy, PYRO_E_INVALID, "NULL argument");
        *strategy = ctx->strategy != 0 ? ctx->strategy : (ctx->tuned_key >= 0 ? ctx->tuned_strategy : 0);
    });
}

int pyro_profile_enable(pyro_ctx* ctx, int on) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->collect_timing();
        ctx->prof = on != 0;
        if (on) {
            std::fill(ctx->prof_ms, ctx->prof_ms + kFamilies, 0.0);
            std::fill(ctx->prof_n, ctx->prof_n + kFamilies, 0);
        }
    });
}

int pyro_profile_read(pyro_ctx* ctx, double* ms, uint64_t* launches) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        ctx->collect_timing();
        for (int f = 0; f < kFamilies; ++f) {
            if (ms) ms[f] = ctx->prof_ms[f];
            if (launches) launches[f] = ctx->prof_n[f];
        }
    });
}

int pyro_mark(pyro_ctx* ctx, int which) {
    return guarded([&] {
        require(ctx != nullptr, PYRO_E_INVALID, "NULL context");
        require(which == 0 || which == 1, PYRO_E_INVALID, "mark must be 0 or 1");
        if (!ctx->marks[which]) HIPCHK(hipEventCreate(&ctx->marks[which]));
        HIPCHK(hipEventRecord(ctx->marks[which], ctx->stream));
    });
}

int pyro_elapsed_ms(pyro_ctx* ctx, double* ms) {
    return guarded([&] {
        require(ctx && ms, PYRO_E_INVALID, "NULL argument");
        require(ctx->marks[0] && ctx->marks[1], PYRO_E_STATE, "both marks must be recorded");
        HIPCHK(hipEventSynchronize(ctx->marks[1]));
        float f = 0.f;
        HIPCHK(hipEventElapsedTime(&f, ctx->marks[0], ctx->marks[1]));
        *ms = f;
    });
}

int pyro_lt_spectrum(const int32_t L[3], double J1, double J2, double J3, double Delta, double* lambda_min,
                     int32_t k_star[3], double* eigvec) {
    return guarded([&] {
        require(L && lambda_min && k_star && eigvec, PYRO_E_INVALID, "NULL argument");
        require(L[0] >= 1 && L[1] >= 1 && L[2] >= 1, PYRO_E_INVALID, "L must be positive");
        require(int64_t(L[0]) * L[1] * L[2] <= (int64_t(1) << 24), PYRO_E_INVALID, "at most 2^24 k points");
        for (double j : {J1, J2, J3, Delta}) require(std::isfinite(j), PYRO_E_INVALID, "couplings must be finite");
        const pyro::LtBondTable table = pyro::pyrochlore_lt_bonds(J1, J2, J3, Delta);
        const int Li[3] = {L[0], L[1], L[2]}, cell[3] = {8, 8, 8};
        const std::string err = pyro::lt_spectrum(table, 16, Li, cell, lambda_min, k_star, eigvec);
        require(err.empty(), PYRO_E_HIP, err);
    });
}

int pyro_batch_create(pyro_ctx* const* ctxs, int n_ctx, pyro_batch** out) {
    return guarded([&] {
        require(ctxs && out && n_ctx >= 1, PYRO_E_INVALID, "a batch needs at least one context");
        *out = nullptr;
        std::set<const pyro_ctx*> seen;
        for (int i = 0; i < n_ctx; ++i) {
            require(ctxs[i] != nullptr, PYRO_E_INVALID, "NULL context");
            ctxs[i]->need_setup();
            require(seen.insert(ctxs[i]).second, PYRO_E_INVALID, "a context appears twice in the batch");
            require(ctxs[i]->device == ctxs[0]->device, PYRO_E_INVALID, "the batch's contexts must share a device");
        }
        auto b = std::make_unique<pyro_batch>();
        b->ctx.assign(ctxs, ctxs + n_ctx);
        HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->ev.resize(3 + 2 * size_t(n_ctx), nullptr);
        for (auto& e : b->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        *out = b.release();
    });
}

int pyro_batch_destroy(pyro_batch* b) {
    return guarded([&] {
        if (!b) return;
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        delete b;
    });
}


int pyro_batch_anneal_steps(pyro_batch* b, int kind, int n_overrelax, int n_sweep, int sample_ssf, int n_steps) {
    return guarded([&] {
        require(b != nullptr, PYRO_E_INVALID, "NULL batch");
        for (pyro_ctx* c : b->ctx) c->anneal_check(kind, n_overrelax, n_sweep, sample_ssf, n_steps);
        if (n_steps == 0) return;
        const int n = int(b->ctx.size());
        const unsigned long long ppr = (unsigned long long)n_sweep * (1 + n_overrelax);
        // the persistent launch takes the eligible members that share the
        // first one's template arguments (precision, shells, per-replica
        // couplings, pass order); the rest sweep on their own
        b->in_launch.assign(size_t(n), 0);
        std::vector<pyro_ctx*> lm;
        if (n_sweep > 0)
            for (int i = 0; i < n; ++i) {
                pyro_ctx* c = b->ctx[i];
                if (!batch_launch_eligible(c)) continue;
                const pyro_ctx* f = lm.empty() ? c : lm[0];
                if (c->prec != f->prec || c->spec_shells != f->spec_shells || c->has_Jrep != f->has_Jrep ||
                    c->fused_order(3, n_overrelax) != f->fused_order(3, n_overrelax))
                    continue;
                lm.push_back(c);
                b->in_launch[i] = 1;
            }
        const bool fused = !lm.empty() && lm[0]->fused_order(3, n_overrelax);
        const bool f64 = !lm.empty() && lm[0]->prec == 64;
        const int shells = lm.empty() ? 3 : lm[0]->spec_shells;
        const bool per_rep = !lm.empty() && lm[0]->has_Jrep;
        std::vector<int32_t> work;
        std::vector<unsigned char> mem;
        if (!lm.empty()) {
            const int cap = f64 ? pyro::pyro_batch_capacity<double>(shells, per_rep, kind, true, fused)
                                : pyro::pyro_batch_capacity<float>(shells, per_rep, kind, true, fused);
            int px = 0, k = 0;
            if (batch_plan(lm, cap, work, px, k)) {
                b->per_xcd = px;
                b->blocks_per_worker = k;
                if (f64) batch_members<double>(b, lm, mem);
                else batch_members<float>(b, lm, mem);
            } else {  // too many replicas to hold at once: every member sweeps on its own
                lm.clear();
                std::fill(b->in_launch.begin(), b->in_launch.end(), 0);
            }
        }
        b->launch_members = int(lm.size());
        // strategies of the members that sweep on their own
        std::vector<int> strat(size_t(n), 0);
        for (int i = 0; i < n; ++i)
            if (!b->in_launch[i]) strat[i] = n_sweep > 0 ? b->ctx[i]->choose_strategy(kind, n_overrelax) : 1;
        for (pyro_ctx* c : b->ctx) c->anneal_upload_counters();
        if (!lm.empty() && (mem != b->h_members || work != b->h_work)) {
            if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
            HIPCHK(hipStreamSynchronize(b->stream));
            if (mem.size() > b->members.bytes) b->members.alloc(mem.size());
            if (work.size() * sizeof(int32_t) > b->work.bytes) b->work.alloc(work.size() * sizeof(int32_t));
            if (size_t(32) * (b->n_items + 8) * sizeof(unsigned int) > b->bar.bytes)
                b->bar.alloc(size_t(32) * (b->n_items + 8) * sizeof(unsigned int));
            HIPCHK(hipMemcpyAsync(b->members.p, mem.data(), mem.size(), hipMemcpyHostToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(b->work.p, work.data(), work.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                  b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
            b->h_members = mem;
            b->h_work = work;
        }
        for (pyro_ctx* c : lm) c->xcd_used = true;
        // the batch's work follows what each member has queued
        hipEvent_t* ev = b->ev.data();
        hipEvent_t* ev_pre = ev + 3;
        hipEvent_t* ev_end = ev + 3 + n;
        for (int i = 0; i < n; ++i) {
            HIPCHK(hipEventRecord(ev_end[i], b->ctx[i]->stream));
            HIPCHK(hipStreamWaitEvent(b->stream, ev_end[i], 0));
        }
        auto step_body = [&]() {
            HIPCHK(hipEventRecord(ev[0], b->stream));
            for (int i = 0; i < n; ++i) {
                pyro_ctx* c = b->ctx[i];
                HIPCHK(hipStreamWaitEvent(c->stream, ev[0], 0));
                c->anneal_set_T();
                if (!b->in_launch[i]) {
                    const bool fe = n_sweep > 0 && c->sweeps_leave_energy(strat[i]);
                    if (n_sweep > 0)
                        c->run_sweeps(strat[i], kind, n_overrelax, n_sweep, c->counters.as<unsigned long long>(), 0ull,
                                      fe ? c->epart.as<double>() : nullptr);
                    c->anneal_observe(sample_ssf, fe, ppr);
                }
                HIPCHK(hipEventRecord(ev_pre[i], c->stream));
            }
            if (!lm.empty()) {
                for (int i = 0; i < n; ++i)
                    if (b->in_launch[i]) HIPCHK(hipStreamWaitEvent(b->stream, ev_pre[i], 0));
                const hipError_t e =
                    f64 ? pyro::launch_pyro_batch<double>(b->members.as<pyro::BatchMember<double>>(), b->work.as<int4>(),
                                                          b->per_xcd, b->n_items, shells, per_rep, kind, n_overrelax,
                                                          n_sweep, true, fused, b->bar.as<unsigned int>(), b->stream)
                        : pyro::launch_pyro_batch<float>(b->members.as<pyro::BatchMember<float>>(), b->work.as<int4>(),
                                                         b->per_xcd, b->n_items, shells, per_rep, kind, n_overrelax,
                                                         n_sweep, true, fused, b->bar.as<unsigned int>(), b->stream);
                require(e != hipErrorNotSupported, PYRO_E_UNSUPPORTED,
                        "the batch launch needs its whole grid resident on the device at once");
                HIPCHK(e);
                HIPCHK(hipEventRecord(ev[1], b->stream));
                for (int i = 0; i < n; ++i) {
                    if (!b->in_launch[i]) continue;
                    pyro_ctx* c = b->ctx[i];
                    HIPCHK(hipStreamWaitEvent(c->stream, ev[1], 0));
                    c->anneal_observe(sample_ssf, true, ppr);
                    HIPCHK(hipEventRecord(ev_end[i], c->stream));
                }
            }
            for (int i = 0; i < n; ++i)
                HIPCHK(hipStreamWaitEvent(b->stream, b->in_launch[i] ? ev_end[i] : ev_pre[i], 0));
        };
        bool use_graph = true;
        std::vector<long long> key = {kind, n_overrelax, n_sweep, sample_ssf};
        for (int i = 0; i < n; ++i) {
            const pyro_ctx* c = b->ctx[i];
            use_graph = use_graph && !c->prof;
            key.insert(key.end(), {(long long)c->graph_gen, strat[i], b->in_launch[i]});
        }
        if (use_graph && (!b->gexec || key != b->g_key)) {
            if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
            b->gexec = capture_graph(b->stream, step_body);
            b->g_key = key;
        }
        for (int s = 0; s < n_steps; ++s) {
            if (use_graph) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
            else step_body();
            for (int i = 0; i < n; ++i) {
                pyro_ctx* c = b->ctx[i];
                if (!c->anneal_advance(sample_ssf, ppr)) continue;
                // a full mu-nu buffer is flushed on the member's stream between steps
                HIPCHK(hipEventRecord(ev[2], b->stream));
                HIPCHK(hipStreamWaitEvent(c->stream, ev[2], 0));
                c->munu_flush();
                HIPCHK(hipEventRecord(ev_end[i], c->stream));
                HIPCHK(hipStreamWaitEvent(b->stream, ev_end[i], 0));
            }
        }
        // what each member does next follows the batch's steps
        HIPCHK(hipEventRecord(ev[2], b->stream));
        for (pyro_ctx* c : b->ctx) {
            HIPCHK(hipStreamWaitEvent(c->stream, ev[2], 0));
            c->h_T.assign(c->R, c->h_sched[c->anneal_step - 1]);
        }
    });
}

int pyro_batch_info(pyro_batch* b, int* launch_members, int* blocks_per_worker, int* workgroups) {
    return guarded([&] {
        require(b != nullptr, PYRO_E_INVALID, "NULL batch");
        if (launch_members) *launch_members = b->launch_members;
        if (blocks_per_worker) *blocks_per_worker = b->launch_members ? b->blocks_per_worker : 0;
        if (workgroups) *workgroups = b->launch_members ? 8 * b->per_xcd : 0;
    });
}

}  // extern "C"
