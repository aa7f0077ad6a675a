/*
 * pyro_cmc.h — C ABI of the MI355X sweep-and-measure library (libpyro_cmc.so).
 *
 * Plain C types only: opaque context, plain pointers and sizes, int status
 * codes.  No exception crosses this boundary; every entry point returns
 * PYRO_OK (0) or a non-zero status, and pyro_last_error() returns the
 * message of the last failure on the calling thread.
 *
 * One context = one lattice geometry + coupling set, R replicas of it in
 * device memory (HBM), one HIP stream.  A context is not thread-safe.
 * Host buffers are caller-owned; device memory is context-owned.
 *
 * Site order everywhere is the reference's storage order: site index
 * i = sl * n_cells + cell, cell = (n0 * L1 + n1) * L2 + n2.
 * Host spin buffers are [N][3] doubles (x, y, z per site), as in the
 * reference's HeisenbergSpin::S.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the pyro_cmc source tree).
 */
#ifndef PYRO_CMC_H
#define PYRO_CMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pyro_ctx pyro_ctx;

/* status codes */
#define PYRO_OK 0
#define PYRO_E_INVALID 1   /* bad argument: maps to std::invalid_argument */
#define PYRO_E_STATE 2     /* call out of order (e.g. sweep before setup) */
#define PYRO_E_LOGIC 3     /* duplicate coupling name: std::logic_error */
#define PYRO_E_HIP 4       /* HIP runtime failure: std::runtime_error */
#define PYRO_E_UNSUPPORTED 5

/* precision of spins, local fields and acceptance; accumulators are FP64 */
#define PYRO_FP32 32
#define PYRO_FP64 64

/* single-site update kinds */
#define PYRO_METROPOLIS 0  /* MC.cpp:111-130  Gaussian-kick Metropolis      */
#define PYRO_LIFTED 1      /* MC.cpp:249-281  lifted (displacement) Metropolis */
#define PYRO_LIFTED_ROT 2  /* MC.cpp:194-244  lifted rotation Metropolis      */
#define PYRO_OVERRELAX 3   /* MC.cpp:145-153  random-angle rotation about h   */

/* Message of the last failed call on this thread ("" if none). */
const char* pyro_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int pyro_version(void);

/* Number of visible HIP devices (0 when none; never fails). */
int pyro_device_count(void);

/* ------------------------------------------------------------------ */
/* Construction.                                                        */
/* Replaces: build_supercell(UnitCellSpecifier, imat33) + MC_runner    */
/* ctor (MC.hpp:103-108), PyroCubicCell (pyrochlore_geometry.hpp:12-32) */
/* Unit cell: orthorhombic with integer edges cell[3]; n_sl sites at    */
/* integer positions sl_pos[n_sl][3]; sl_kind[s] is the reference's     */
/* pyro_sl (which displacement list of a coupling applies, and which    */
/* endpoint of a bond is the source).  Supercell: L[0] x L[1] x L[2].   */
/* precision: PYRO_FP32 or PYRO_FP64.  seed keys the counter-based RNG. */
/* ------------------------------------------------------------------ */
int pyro_create(int n_sl, const int32_t* sl_pos, const int32_t* sl_kind,
                const int32_t cell[3], const int32_t L[3], int n_replicas,
                int precision, uint64_t seed, pyro_ctx** out);
int pyro_destroy(pyro_ctx* ctx);

/* MC_runner::define_coupling (MC.hpp:122-125, MC.cpp:56-66).
 * rel_vecs holds, for kind k = 0..n_kinds-1, counts[k] displacement
 * vectors (int32 triples) back to back.  J is a row-major 3x3 matrix
 * applied by the source endpoint (lower pyro_sl); the target endpoint
 * applies J^T.  Duplicate names give PYRO_E_LOGIC. */
int pyro_define_coupling(pyro_ctx* ctx, const char* name, int n_kinds,
                         const int32_t* counts, const int32_t* rel_vecs,
                         const double J[9]);

/* Convenience: the nine couplings of build_J1J2J3_h (cli_bits.hpp:144-207):
 * J1 as six sublattice-pair XXZ specs J1*[I + (Delta-1) z_mu z_nu^T], then
 * J2, J3a, J3b isotropic.  Requires the 16-site pyrochlore cubic cell. */
int pyro_define_J1J2J3(pyro_ctx* ctx, double J1, double J2, double J3,
                       double Delta);

/* Pyrochlore neighbour displacement tables by kind p = 0..3
 * (pyrochlore_geometry.hpp:110-149): which = 1 nn1 (|d|^2 = 8), 2 nn2 (24),
 * 3 nn3a (32, bond midpoint on a site), 4 nn3b (32, midpoint empty).
 * counts[4] receives the list lengths, vecs up to cap int32 triples.
 * Host-only; needs no device. */
int pyro_pyrochlore_shell(int which, int32_t counts[4], int32_t* vecs, int cap);

/* MC_runner::setup_lattice (MC.cpp:15-52): builds the int32 neighbour
 * table, per-slot coupling table and the graph colouring, uploads them. */
int pyro_setup_lattice(pyro_ctx* ctx);

/* Geometry / topology queries.
 * info[0]=N sites, [1]=n_cells, [2]=n_sl, [3]=replicas, [4]=precision,
 * [5]=neighbour slots per site (after dropping zero couplings),
 * [6]=number of colours, [7]=1 if every coupling is isotropic. */
int pyro_info(pyro_ctx* ctx, int64_t info[8]);
/* Site positions, [N][3] int32 (the reference's HeisenbergSpin::ipos). */
int pyro_site_positions(pyro_ctx* ctx, int32_t* pos);
/* Neighbour table [N][n_slot] and its per-slot 3x3 couplings [N][n_slot][9];
 * padded slots hold the site itself with a zero matrix. */
int pyro_neighbours(pyro_ctx* ctx, int32_t* nbr, double* J);
/* colour_of_site[N]; sites of one colour share no bond. */
int pyro_colouring(pyro_ctx* ctx, int32_t* colour_of_site);

/* ------------------------------------------------------------------ */
/* State.                                                               */
/* ------------------------------------------------------------------ */
/* MC_runner::set_global_field (MC.cpp:69-71); replica -1 = all. */
int pyro_set_field(pyro_ctx* ctx, int replica, const double B[3]);
/* MC_parameters::T_ref (MC.hpp:79-82). */
int pyro_set_T_ref(pyro_ctx* ctx, double T_ref);
/* Temperature of each replica (T[R]).  Also resets the tempering
 * permutation to the identity (replica r sits in slot r). */
int pyro_set_temperatures(pyro_ctx* ctx, const double* T);
/* Spin upload/download for one replica, host [N][3] doubles
 * (HeisenbergSpin::S; save_spin_state layout, MC.cpp:371-406). */
int pyro_set_spins(pyro_ctx* ctx, int replica, const double* S);
int pyro_get_spins(pyro_ctx* ctx, int replica, double* S);
/* Lifted directions, int8 +-1 per site (HeisenbergSpin::lifted_dir). */
int pyro_set_lifted_dir(pyro_ctx* ctx, int replica, const int8_t* dir);
int pyro_get_lifted_dir(pyro_ctx* ctx, int replica, int8_t* dir);
/* Uniform-on-sphere spins for every replica from `seed` (counter RNG);
 * lifted directions reset to +1. */
int pyro_randomize_spins(pyro_ctx* ctx, uint64_t seed);
/* Number of single-site passes done so far (the RNG pass counter). */
int pyro_pass_count(pyro_ctx* ctx, uint64_t* passes);

/* ------------------------------------------------------------------ */
/* Sweeps.                                                              */
/* MC_runner::sweep_local_Metropolis / sweep_lifted_Metropolis /        */
/* sweep_lifted_Metropolis_rot (MC.cpp:155-163, 283-303): one pass of   */
/* `kind` over every site, then n_overrelax overrelaxation passes,      */
/* repeated n_sweeps times, for all replicas at their temperatures.     */
/* accepted[R] (may be NULL) receives the accepted-update count of the  */
/* first pass kind summed over the n_sweeps sweeps.                      */
/* ------------------------------------------------------------------ */
int pyro_sweep(pyro_ctx* ctx, int kind, int n_overrelax, int n_sweeps,
               uint64_t* accepted);
/* MC_runner::overrelax_all (MC.cpp:145-153), n_passes times. */
int pyro_overrelax(pyro_ctx* ctx, int n_passes);

/* ------------------------------------------------------------------ */
/* Observables on the current configuration.                            */
/* ------------------------------------------------------------------ */
/* MC_runner::local_field (MC.cpp:87-96) for every site of one replica,
 * host [N][3] doubles; the global field is NOT subtracted. */
int pyro_local_fields(pyro_ctx* ctx, int replica, double* h);
/* MC_runner::total_energy_per_unit_cell (MC.cpp:305-312) per replica,
 * e[R], summed in FP64. */
int pyro_energy_per_cell(pyro_ctx* ctx, double* e);

/* Allocation-free variant of pyro_energy_per_cell.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Needs no context and calls no hipMalloc / hipFree.
 *   spins: device memory, [n_replicas][16 L0 L1 L2][3] FP32 or FP64 as
 *          `precision` says, in the context's site order (the pyrochlore
 *          cubic cell); read only.
 *   J1 (non-zero), J2, J3, Delta: the couplings of pyro_define_J1J2J3,
 *   B[3] (host): the global field of every replica.
 *   e: device memory, double[n_replicas], receives the energy per cell.
 *   stream: a hipStream_t (NULL: the null stream).  The call only enqueues:
 *          e is ready once the stream reaches that point.
 * Results equal pyro_energy_per_cell's on a context holding the same
 * spins, couplings and field, bit for bit (same kernels).  The partial sums
 * live in one 4 MiB device array per GPU, held until the process ends:
 * n_replicas x (energy blocks per replica, L0 L1 L2 / 4 or fewer) + 3
 * n_replicas doubles must fit in it (else PYRO_E_INVALID).  Calls that
 * overlap on one GPU (two streams, or two threads) share that array and
 * race; serialise them. */
int pyro_energy_per_cell_noalloc(const void* spins, int precision, const int32_t L[3], int n_replicas,
                                 double J1, double J2, double J3, double Delta, const double B[3],
                                 double* e, void* stream);

/* ------------------------------------------------------------------ */
/* Replica exchange (temper.cpp:418-474 without MPI).                   */
/* T_ladder[R]: slot temperatures, slot 0 hottest (descending, like the */
/* reference's T_grid / rank order).  Replica r starts in slot r.       */
/* ------------------------------------------------------------------ */
int pyro_pt_init(pyro_ctx* ctx, const double* T_ladder, uint64_t swap_seed);
/* One tempering round: n_sweep sweeps (kind + n_overrelax OR passes) of
 * every replica at its slot temperature; if sample_ssf, one SSF sample of
 * every replica into its slot's accumulator (sample_ssf = 1: after the
 * sweeps, as temper.cpp:424-429; sample_ssf = 2: after each of the
 * sweeps); the energy per cell of every replica is added to its slot's E,
 * E^2, n; then the checkerboard exchange of slot pairs (s, s+1) with
 * s = round parity (mod 2).  n_rounds rounds are run back to back
 * (captured once as a HIP graph). */
int pyro_pt_rounds(pyro_ctx* ctx, int kind, int n_overrelax, int n_sweep,
                   int sample_ssf, int n_rounds);
/* Per-pair swap statistics attempts[R-1], accepts[R-1] (pair s <-> s+1),
 * and per-slot energy sums E[R], E2[R], n[R] (temper.cpp:432-435,
 * 478-508).  Any pointer may be NULL. */
int pyro_pt_stats(pyro_ctx* ctx, uint64_t* attempts, uint64_t* accepts,
                  double* E, double* E2, uint64_t* n);
/* replica_of_slot[R]. */
int pyro_pt_permutation(pyro_ctx* ctx, int32_t* replica_of_slot);
/* Zero the swap statistics and energy sums (burn-in -> sampling). */
int pyro_pt_reset_stats(pyro_ctx* ctx);
/* Exchange after every `every`-th round only (default 1, the reference's
 * temper.cpp:418-474 cadence): the rounds in between still sweep, sample
 * the SSF and the energies, so a schedule with samples every sweep and
 * exchanges every k sweeps is pyro_pt_rounds with n_sweep = 1 and this
 * interval k.  Exchange x (x = 0, 1, ... at the end of rounds k - 1,
 * 2k - 1, ...) has parity x mod 2 and its own random numbers; with k = 1
 * the exchanges are those without the setting.  No counterpart in the
 * reference, whose rounds always end with an exchange. */
int pyro_pt_set_exchange_interval(pyro_ctx* ctx, int every);
/* Accumulate the current energy per cell of every replica into the
 * energy sums of its slot (energy_manager::sample, energy_manager.hpp:38-44). */
int pyro_energy_sample(pyro_ctx* ctx);

/* ------------------------------------------------------------------ */
/* Batches of independent anneals.  No counterpart in the reference,    */
/* which runs one anneal per process (seeds merged afterwards,          */
/* acc_runs.py:13-57); here every replica of a context can be its own   */
/* model and follow a device-side temperature schedule.                 */
/* ------------------------------------------------------------------ */
/* Per-replica pyrochlore couplings J1J2J3Delta[R][4] = (J1, J2, J3,
 * Delta) for a context whose couplings come from pyro_define_J1J2J3; a
 * replica may not switch on a shell the context defined as zero (the
 * colouring is the context's).  Fields B stay per replica
 * (pyro_set_field).  NULL returns to the context's couplings.  Needs the
 * pyrochlore strategies (0 = automatic, 3, 4, 5). */
int pyro_set_replica_couplings(pyro_ctx* ctx, const double* J1J2J3Delta);
/* Temperature schedule T[n_steps] shared by all replicas (the anneal
 * grid of anneal.cpp:164-227 / cli_bits.hpp:250-289); resets the step
 * counter, the energy history and the permutation (slot r = replica r). */
int pyro_anneal_schedule(pyro_ctx* ctx, const double* T, int n_steps);
/* n_steps anneal steps, each: every replica at the schedule's next
 * temperature, n_sweep sweeps (kind + n_overrelax OR passes), if
 * sample_ssf one SSF sample of every replica into slot r (or the slot
 * pyro_pt_ssf_map gives replica r), then its energy
 * per cell recorded in the history (energy_manager::sample per step,
 * anneal.cpp:203-206).  Captured once as a HIP graph and only enqueued:
 * contexts (each with its own stream) run concurrently until synchronised
 * by a reading call. */
int pyro_anneal_steps(pyro_ctx* ctx, int kind, int n_overrelax, int n_sweep,
                      int sample_ssf, int n_steps);
/* Energy per cell of every replica at every schedule step, E[R][n_steps]
 * (zero for steps not yet run). */
int pyro_anneal_history(pyro_ctx* ctx, double* E);

/* A batch of contexts stepped together, e.g. one context per lattice size
 * of an uneven batch of anneals.  No counterpart in the reference, which
 * runs one anneal per process and merges them afterwards
 * (acc_runs.py:13-57).  The contexts stay owned by the caller and must
 * outlive the batch. */
typedef struct pyro_batch pyro_batch;
int pyro_batch_create(pyro_ctx* const* ctxs, int n_ctx, pyro_batch** out);
/* pyro_anneal_steps(ctx, ...) for every context of the batch, with the
 * same results bit for bit, as one HIP graph per step: the contexts whose
 * sweeps take the per-colour stencil strategies under the automatic choice,
 * or that are forced to strategy 6, sweep together in one persistent launch (every replica of every such
 * context, whole sweeps, each replica's workgroups on one XCD, stage waits
 * bounded as set by pyro_set_wait_timeout), the others (the tiled strategy
 * on large lattices, forced strategies) sweep on their own in parallel
 * branches of the same graph.  Every context needs an anneal schedule with
 * n_steps steps left.  Only enqueued: reading calls on a context wait for
 * the batch's steps. */
int pyro_batch_anneal_steps(pyro_batch* b, int kind, int n_overrelax, int n_sweep, int sample_ssf,
                            int n_steps);
/* After pyro_batch_anneal_steps: the contexts in the persistent launch,
 * the cell blocks each of its workgroups takes per colour stage, and its
 * workgroups (0, 0, 0 without one). */
int pyro_batch_info(pyro_batch* b, int* launch_members, int* blocks_per_worker, int* workgroups);
int pyro_batch_destroy(pyro_batch* b);

/* ------------------------------------------------------------------ */
/* Static structure factor (ssf_manager.hpp:42-57, 260-293).            */
/* ------------------------------------------------------------------ */
/* n_corr correlators (alpha[c], beta[c] in {0,1,2}); n_slots accumulator
 * slots (temperatures).  store_error_term keeps the (Re S)^2, (Im S)^2
 * sums (static_corr_2).  munu != 0 also accumulates the sublattice-
 * resolved C^{ab}_{mu nu}(K) (temper.cpp:98-159 layout) through buffered
 * FP32-MFMA products flushed into FP64; munu_buffer = samples per flush. */
int pyro_ssf_configure(pyro_ctx* ctx, int n_corr, const int32_t* alpha,
                       const int32_t* beta, int n_slots, int store_error_term,
                       int munu, int munu_buffer);
/* Accumulator slot of each tempering slot for the SSF samples taken in
 * pyro_pt_rounds: map[s] in [-1, n_slots), -1 = that rung is not sampled
 * (the reference samples only its T_sample rungs, temper.cpp:426-429).
 * No two rungs may share a slot.  NULL clears the map (then tempering slot
 * s accumulates into slot s).  Anneal steps (pyro_anneal_steps, where slot
 * r is replica r) follow the same map, so a schedule with several sampling
 * temperatures (ssf_manager::new_T per T_sample, anneal.cpp:208-221) sends
 * each temperature's samples to its own slots by changing the map between
 * calls; a new map takes effect without re-capturing the step graph. */
int pyro_pt_ssf_map(pyro_ctx* ctx, const int32_t* map);
/* One SSF sample of every replica r with slot_of_replica[r] >= 0 into that
 * slot.  slot_of_replica == NULL: use each replica's tempering slot. */
int pyro_ssf_sample(pyro_ctx* ctx, const int32_t* slot_of_replica);
/* Read slot accumulators: S[n_corr][n_k][2] and S2[n_corr][n_k][2]
 * (either may be NULL), n_samples.  k order: (K0*L1 + K1)*L2 + K2. */
int pyro_ssf_read(pyro_ctx* ctx, int slot, double* S, double* S2,
                  uint64_t* n_samples);
/* Read sublattice-resolved accumulators of one slot for correlator c:
 * C[n_k][n_sl][n_sl][2] (requires munu). */
int pyro_ssf_read_munu(pyro_ctx* ctx, int slot, int c, double* C);
/* Zero one slot's accumulators (slot -1 = all). */
int pyro_ssf_reset(pyro_ctx* ctx, int slot);
/* Raw per-sublattice DFT of one replica's spins, ft[3][n_sl][n_k][2]
 * (save_ft_spin_state layout, MC.cpp:320-368); needs no SSF configuration. */
int pyro_spin_ft(pyro_ctx* ctx, int replica, double* ft);

/* ------------------------------------------------------------------ */
/* Execution control and timing.                                        */
/* ------------------------------------------------------------------ */
int pyro_synchronize(pyro_ctx* ctx);
/* Sweep launch strategy: 0 = automatic (default), 1 = one launch per
 * colour (neighbour table), 2 = one resident workgroup per replica looping
 * over all colours of a sweep (neighbour table), 3 = one launch per colour
 * with the compile-time pyrochlore stencil, 4 = resident workgroup per
 * replica with the stencil (at most 2048 cells), 5 = tiled sweep (tiles and
 * their halo in LDS, 8 checkerboard phases; lattices a tile shape divides),
 * 6 = XCD-resident sweep (the per-colour work of 3 for whole sweeps in one
 * launch, every workgroup of a replica on one XCD, the colour stages
 * separated by waits on counters in that XCD's L2; when the grid fits on
 * the device at once).  3 to 6 apply when the couplings were defined by
 * pyro_define_J1J2J3 alone.  Strategies 1 and 2 give bit-identical results,
 * as do 3, 4 and 6 (same colour order, per-site counters and arithmetic);
 * the automatic choice takes 5 on lattices of at least 4096 cells that a
 * tile shape divides, and otherwise times the members of its bit-identical
 * set and keeps the fastest, so it never changes results.  Sweeps report
 * PYRO_E_UNSUPPORTED for a forced strategy that does not apply. */
int pyro_set_strategy(pyro_ctx* ctx, int strategy);
/* Bound, in seconds, on each wait between the colour stages of the
 * XCD-resident sweep (strategy 6) and of batch launches (pyro_batch_*;
 * default 0.25 s), counted from the workgroup's own arrival at the end of
 * the previous stage.  A wait past it ends the launch, and the next call that reads results reports PYRO_E_HIP ("a stage
 * wait timed out") instead of returning them.  No counterpart in the
 * reference (its sweeps have no waits). */
int pyro_set_wait_timeout(pyro_ctx* ctx, double seconds);
int pyro_get_strategy(pyro_ctx* ctx, int* strategy);
/* Tile shape of the tiled sweep (strategy 5), in cells per axis: one of
 * 4x4x4, 6x6x4, 8x4x4, 4x4x8, 8x8x4, applicable when each L_a is divisible
 * by twice the edge.  T = NULL restores the automatic choice: the first of
 * 6x6x4, 8x8x4, 4x4x4, 4x4x8, 8x4x4 that applies and gives at least 8
 * tiles per checkerboard phase, else the first that applies.  The tile shape
 * changes the tiled visiting order (pyro_sweep_order reports it), so it
 * is a fixed rule, never a timing choice.  pyro_get_tile_shape returns the
 * shape in use (0, 0, 0 when none applies). */
int pyro_set_tile_shape(pyro_ctx* ctx, const int32_t T[3]);
/* Pass order of the pyrochlore stencil sweeps (strategies 3, 4 and 5; the
 * generic strategies 1 and 2 always run pass by pass; 3 and 4 stay
 * bit-identical in either order).  0 (default): pass by pass, the order
 * the reference's sweeps follow (MC.cpp:155-163, 283-303: every site's
 * first-kind update, then every site's OR update).  1 (fused): within each
 * colour stage every site takes its first-kind update and then its first
 * OR update back to back, from one evaluation of its local field (no
 * bonded site changes in between); OR passes 2..n_overrelax follow as
 * passes of their own (two OR rotations about one field would make one).
 * Its uniforms for the sweep are drawn in order from its Philox blocks:
 * uniform j is component j % 5 of the block with pass index p0 + j / 5 (p0
 * the sweep's first pass), the first kind taking 5 (Metropolis), 2
 * (lifted) or 1 (lifted rotation) and OR pass k the next one, so a lifted
 * + OR sweep draws one block per site and a Metropolis + 2 OR sweep two.
 * Per-colour launches (strategy 3) then run 16 launches per sweep for the
 * first kind with one OR pass, and 16 more per further OR pass.  Both
 * orders are sequential single-site schedules of the same updates; the
 * order changes the Markov chain, not its stationary distribution, and
 * pyro_sweep_order reports it.  No counterpart in the reference. */
int pyro_set_fused_passes(pyro_ctx* ctx, int on);
int pyro_get_tile_shape(pyro_ctx* ctx, int32_t T[3]);
/* The site visiting order of one sweep under the strategy the next sweep
 * will use and the pass order (pyro_set_fused_passes): site[k] and
 * pass_offset[k] (0 = the sweep's first pass kind, 1..n_overrelax = OR
 * passes) for k < N * (1 + n_overrelax).  Sites with
 * no bond between them may run in any relative order, so this is one of
 * the sequential orders the parallel sweep equals. */
int pyro_sweep_order(pyro_ctx* ctx, int n_overrelax, int32_t* site, int32_t* pass_offset);

/* Kernel timing with HIP events on the context's stream.  When enabled,
 * every launch is bracketed by events and its duration is added to the
 * counter of its kernel family:
 *   0 = sweep passes, 1 = energy, 2 = exchange, 3 = SSF transforms,
 *   4 = SSF contraction (with the flush of buffered contracted samples),
 *   5 = sublattice-resolved staging of samples,
 *   6 = sublattice-resolved MFMA products flushed into the accumulators.
 * ms[7], launches[7]. */
int pyro_profile_enable(pyro_ctx* ctx, int on);
int pyro_profile_read(pyro_ctx* ctx, double* ms, uint64_t* launches);
/* Elapsed milliseconds between two marks recorded on the context stream. */
int pyro_mark(pyro_ctx* ctx, int which);
int pyro_elapsed_ms(pyro_ctx* ctx, double* ms);

/* ------------------------------------------------------------------ */
/* Luttinger-Tisza spectrum (luttinger_tisza.cpp:31-181).               */
/* ------------------------------------------------------------------ */
/* For the pyrochlore J1-J2-J3 XXZ model on the L[0] x L[1] x L[2] grid of
 * k (cubic cell of edge 8, q_a = 2 pi K'_a / (8 L_a) with the centred
 * K' = K < L/2 ? K : K - L) builds the 48 x 48 Hermitian
 * M(k) = sum over directed bonds alpha -> beta of J e^{i q.d} in block
 * (alpha, beta) and returns its lowest eigenvalue lambda_min[k] for every
 * k (flat order (K0 L1 + K1) L2 + K2).  k_star[3] receives the first k
 * (in that order) of the strict minimum and eigvec[48][2] the (re, im)
 * eigenvector of the lowest eigenvalue there, rows ordered spin x, y, z of
 * sublattice 0, 1, ...  The LT energy per spin is lambda_min / 2.  Uses
 * the current device; no context is needed.  Replaces the Eigen loop of
 * luttinger_tisza.cpp:147-181. */
int pyro_lt_spectrum(const int32_t L[3], double J1, double J2, double J3, double Delta,
                     double* lambda_min, int32_t k_star[3], double* eigvec);

#ifdef __cplusplus
}
#endif

#endif /* PYRO_CMC_H */
