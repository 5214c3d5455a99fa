/*
 * bafqmc.h -- C ABI of the MI355X solver for the BAFQMC Markov-chain sweep.
 *
 * The reference (YinkaiYu/BAFQMC) has no library API: its sweep is reached by
 * running the executable `mpirun -np R bosonDQMC.out` in a run directory
 * (src/number_conserving/run_paper.py:113,162; src/pairing/run_paper.py:100-101),
 * one MPI rank per Markov chain.  Each entry point below replaces one step of
 * that executable's main program (src/number_conserving/src/main.f90:22-83,
 * identical structure in src/pairing/src/main.f90) for a whole batch of
 * chains at once; the Python host (bafqmc_amd) keeps the file contract.
 *
 * Conventions: plain pointers and sizes, host memory on every pointer
 * argument, row-major matrices, complex numbers as interleaved (re, im)
 * doubles.  Every call returns 0 on success and a nonzero code otherwise, with
 * a message from bafq_last_error().  One handle = one batch of chains on one
 * GPU stream; calls on one handle must not overlap; distinct handles are
 * independent.
 */
#ifndef BAFQMC_H
#define BAFQMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bafq_handle bafq_handle;

enum { BAFQ_NUMBER_CONSERVING = 0, BAFQ_PAIRING = 1 };

/* One Hamiltonian + discretisation shared by a group of chains.
 * Replaces read_input/Params_set (src/number_conserving/src/calc_basic.f90:48-106;
 * pairing adds RDelta, src/pairing/src/calc_basic.f90:51) and the kinetic
 * propagators of opT_set (src/number_conserving/src/non_interact.f90:74-109;
 * src/pairing/src/non_interact.f90:121-178).  The solver works in the reduced
 * form: N = Nlx*Nly (number conserving) or N = 2*Nlx*Nly (pairing, basis
 * b_1..b_Ns, c+_1..c+_Ns).  expk_p / expk_m are the real N x N matrices
 * exp(-dtau A) and exp(+dtau A) in that basis, built by the host. */
typedef struct bafq_model_desc {
    int32_t kind;
    int32_t nlx, nly, ltrot, nwrap;
    double beta, u1, u2, mu, rdelta, shift_loc;
    const double* expk_p;
    const double* expk_m;
} bafq_model_desc;

/* Model_init + Prop%make + WrList%make + Stabilize_init for every chain
 * (src/number_conserving/src/main.f90:29-35).  chain_model[c] selects the
 * model of chain c. */
int bafq_create(const bafq_model_desc* models, int n_models, const int32_t* chain_model,
                int n_chains, bafq_handle** out);
void bafq_destroy(bafq_handle* h);
const char* bafq_last_error(const bafq_handle* h);

/* Execution schedule of the chains of one model (no counterpart in the
 * reference, which runs one chain per MPI rank with LAPACK/BLAS calls):
 * path  -1 choose, 0 Green's function and work matrices in device memory,
 *       1 G and the stabilisation work matrix resident in LDS (N <= 96);
 * delayed  -1 widest that fits, 0 per-site rank-1/rank-2 updates, or 4, 8,
 *       12, 16 accepted updates buffered in LDS before one MFMA flush into G
 *       (either path).
 * Every schedule computes the same Markov chain (identical up to rounding);
 * bafqmc_amd.autotune picks one by timing.  bafq_create uses {-1, -1}. */
typedef struct bafq_schedule {
    int32_t path;
    int32_t delayed;
} bafq_schedule;

/* bafq_create with a schedule per model (schedule may be NULL).  A schedule
 * that cannot run (LDS path for a model that does not fit) is an error. */
int bafq_create_scheduled(const bafq_model_desc* models, int n_models, const int32_t* chain_model,
                          int n_chains, const bafq_schedule* schedule, bafq_handle** out);
/* The schedule in effect for a model. */
int bafq_get_schedule(const bafq_handle* h, int model, bafq_schedule* out);
/* Pairs of workgroups that share a chain in the device-memory-path kernels
 * of bafq_sweep, bafq_prepare and bafq_measure_pass (0: one workgroup per
 * chain).  No counterpart in the reference (one chain per MPI rank).
 * bafq_create pairs the workgroups when every device-memory-path chain then
 * still has its own resident pair and every model on that path uses
 * submatrix updates (delayed > 0); with LDS-path chains in the batch, when
 * both paths' kernels hold the same workgroups per CU and the pairs leave at
 * least one workgroup for the LDS path, whose grid is then cut to fit beside
 * them.  The environment variable BAFQ_PAIRS=0 turns this off, BAFQ_PAIRS=1
 * pairs as many workgroups as are resident at once (batches without LDS-path
 * chains).  Results are identical to the unpaired kernels', bit for bit. */
int bafq_pairs(const bafq_handle* h);

/* Sizes: matrix dimension N of a chain, doubles per chain in bafq_read_bin's
 * obs block, doubles per chain in its stats block, fields per chain. */
int bafq_dims(const bafq_handle* h, int chain, int32_t* n, int32_t* obs_stride,
              int32_t* stats_stride, int32_t* n_fields);

/* conf_in / conf_out (src/number_conserving/src/fields.f90:35-113,236-278):
 * fields phi(ns, site, nt) in Fortran order, i.e. [nt][site][ns] per chain,
 * concatenated over chains; one RNG seed per chain. */
int bafq_set_fields(bafq_handle* h, const double* phi, const int32_t* seeds);
int bafq_get_fields(bafq_handle* h, double* phi, int32_t* seeds);

/* n calls of Local_sweep_therm (src/number_conserving/src/local_sweep.f90:64-78). */
int bafq_warmup(bafq_handle* h, int n);

/* Local_sweep_pre (local_sweep.f90:80-94): B(beta,0) checkpoints and G(beta). */
int bafq_prepare(bafq_handle* h);

/* n_sweeps x (sweep_L; sweep_R) = Local_sweep with Nsweep = n_sweeps
 * (local_sweep.f90:136-161); measures every slice when measure != 0.
 * Chains are taken from a work queue in decreasing order of cost by a
 * persistent grid (one launch per path; the launches of a (n_sweeps, measure)
 * pair are captured once in a HIP graph and replayed). */
int bafq_sweep(bafq_handle* h, int n_sweeps, int measure);

/* Conventions of the time-displaced pass. */
enum { BAFQ_TAU_BOSONIC = 1, BAFQ_TAU_REFERENCE = 2 };

/* bafq_sweep with the time-displaced pass before every sweep_R: Local_sweep
 * with toggle = .true. (src/number_conserving/src/local_sweep.f90:119-123),
 * i.e. Dyn_reset + Dyn_sweep_R (dynamics.f90:38-57) on the fields sweep_L
 * left: G(tau,0), G(0,tau), G(tau,tau) propagated slice by slice from G(0)
 * (multiply.f90:48-64,88-97) and recomputed every Nwrap slices by Wrap_tau /
 * stab_green_big (stabilization.f90:192-254,368-404), whose drifts go to the
 * stats slots 20-22 (running maxima of G(0,tau), G(tau,0), G(tau,tau), the
 * reference's "Max diff GRtt/GRt0/GR0t" in that order), 23-25 (sums, reset by
 * bafq_read_bin) and 26 (count).  The pass touches no field, seed, G or
 * equal-time accumulator: the chain is the one bafq_sweep runs.
 * mode BAFQ_TAU_BOSONIC: the 2N x 2N block matrix with -D_L, whose inverse
 * holds the bosonic chain's functions (G(0) = (1 - B)^{-1}); mode
 * BAFQ_TAU_REFERENCE: the reference's code as written, +D_L
 * (stabilization.f90:206-208, the fermionic form) and the U2 operator
 * propagated with the U1 field (dynamics.f90:50-52), whose drift lines it
 * reproduces.  record != 0 keeps (G00, Gt0, G0t, Gtt) at tau = 0..Ltrot-1 of
 * the last sweep's pass (Obs_tau_calc's arguments, obser_tau.f90:39-54) for
 * bafq_get_green_tau.  The workspace (17 N^2 complex per chain, plus
 * 4 Ltrot N^2 when recording) is allocated on first use. */
int bafq_sweep_tau(bafq_handle* h, int n_sweeps, int measure, int mode, int record);

/* The recorded time-displaced functions of one chain: Ltrot x 4 x N x N
 * complex, [tau][G00, Gt0, G0t, Gtt] (PropGreen, process_matrix.f90:17-24). */
int bafq_get_green_tau(bafq_handle* h, int chain, double* out);

/* n_passes measurement passes: one left sweep each with the fields frozen (no
 * proposals; the RNG streams are untouched) that measures the equal-time
 * estimators at every slice into the same accumulators as bafq_sweep, as
 * sweep_L measures them (src/number_conserving/src/local_sweep.f90:96-111).
 * mode 0: every slice; mode 1: only at the Nst stabilisation points nt =
 * Ltrot, Ltrot - Nwrap, ..., Nwrap, where Wrap_L computes G afresh, without
 * the G wraps between them; mode 2: the current Green's function once
 * (G(beta) after bafq_prepare or a move; no pass, n_passes ignored, nothing
 * but the accumulators changes).  Every slice's measurement estimates the same
 * equal-time averages (the trace is cyclic in imaginary time).  No
 * counterpart in the reference, whose only measurement is inside Local_sweep:
 * this measures a configuration reached by a global move
 * (bafq_global_propose / bafq_hmc_propose) without a local sweep.  G must be
 * current (after bafq_prepare, a sweep or a move). */
int bafq_measure_pass(bafq_handle* h, int n_passes, int mode);

/* Per-chain observable sums, sample counts and counters accumulated since the
 * previous call, then reset (replaces Obs_equal%ave + Fourier%preq's per-rank
 * values, src/number_conserving/src/fourier_trans.f90:350-457; acceptance
 * counters of AccCounter, operator_Hubbard.f90:104-129; the drift of
 * compare_mat, stabilization.f90:256-266).  obs: n_chains * obs_stride,
 * stats: n_chains * stats_stride. */
int bafq_read_bin(bafq_handle* h, double* obs, double* stats);

/* The per-chain counters of bafq_read_bin without reading or resetting
 * anything; slot 7 holds log|det G| of the chain's current fields (from its
 * last stabilised Green's function), the quantity the exact acceptance of a
 * global move needs (the weight of src/number_conserving/src/pole_diagnostics.f90:130-131).
 * No counterpart in the reference's executable, which runs no global moves. */
int bafq_get_stats(bafq_handle* h, double* stats);

/* Current equal-time Green's function of one chain, N x N complex
 * (Prop%Gr, src/number_conserving/src/process_matrix.f90:8-10). */
int bafq_get_green(bafq_handle* h, int chain, double* g);

/* Device time in milliseconds of the most recent warmup/prepare/sweep kernel,
 * measured with HIP events on the stream the kernel ran on (replaces the
 * SYSTEM_CLOCK wall time of src/number_conserving/src/main.f90:26-27,67-75). */
int bafq_last_kernel_ms(const bafq_handle* h, double* ms);

/* Self-test hooks with no counterpart in the reference: run the device
 * complex GEMM C = op(A) op(B) (op 0: none, 1: conjugate transpose) and the
 * column-pivoted QR on host matrices, for the numerics tests.  QR variant 0 is
 * the workgroup-cooperative routine on a matrix in device memory, variant 1
 * the column-parallel routine on a matrix in LDS (n <= 96). */
int bafq_debug_gemm(int n, const double* a, int op_a, const double* b, int op_b, double* c);
int bafq_debug_qrp(int n, double* a, double* tau, int32_t* perm, double* d, int variant);
/* Lane layout of the FP64 MFMA 16x16x4 tile operation: out[p][lane][r] =
 * accumulator r of `lane` for A = one-hot at lane p, B = 1000 + lane
 * (64 x 64 x 4 doubles). */
int bafq_debug_mfma_probe(double* out);
/* The same for the FP32 MFMA 16x16x4 used by the network's convolutions. */
int bafq_debug_mfma32_probe(float* out);
/* Cross-lane operations of one 64-lane wavefront on in[64] / idx[64]:
 * out[0:64] sums over aligned 16-lane rows, out[64:128] sums over aligned
 * quads, out[128:192] lane 37's value in every lane, out[192:256] and
 * out[256:320] the wavefront maximum of in and the lowest idx attaining it. */
int bafq_debug_lane_ops(const double* in, const int32_t* idx, double* out);

/* ---------------------------------------------------------------------
 * Learned global proposals: a periodic 3D convolutional network over HS
 * fields, trained and evaluated on the GPU (no counterpart in the reference).
 * Fields go in and out as float32 [batch][channels][lt][ly][lx]; parameters
 * are the concatenation over layers of W[cout][cin][3][3][3] and b[cout];
 * hidden layers use ReLU, the last layer is linear.  The loss is the mean
 * squared error against a target of the output's shape.
 * --------------------------------------------------------------------- */
typedef struct bafq_net bafq_net;
typedef struct bafq_net_desc {
    int32_t lt, ly, lx;       /* lattice: time slices, rows, columns (periodic) */
    int32_t channels;         /* field channels in and out */
    int32_t hidden;           /* channels of the hidden layers */
    int32_t layers;           /* convolution layers, >= 2 */
} bafq_net_desc;

int bafq_net_create(const bafq_net_desc* d, int max_batch, bafq_net** out);
void bafq_net_destroy(bafq_net* h);
const char* bafq_net_last_error(const bafq_net* h);
int64_t bafq_net_param_count(const bafq_net* h);
/* set (resets the Adam state) / get the parameters */
int bafq_net_set_params(bafq_net* h, const float* params);
int bafq_net_get_params(bafq_net* h, float* params);
/* y = f(x) */
int bafq_net_forward(bafq_net* h, int batch, const float* x, float* y);
/* loss = |f(x) - target|^2 / size and its gradient (grad may be NULL) */
int bafq_net_gradient(bafq_net* h, int batch, const float* x, const float* target, double* loss, float* grad);
/* y = f(x) without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  params, x and y are DEVICE
 * pointers the caller owns (params in the layout of bafq_net_set_params, x
 * and y [batch][channels][lt][ly][lx], x and y not overlapping).  The hidden
 * activations live in the library's static device scratch of 16 Mi floats
 * (64 MiB per GPU, held until the process ends); the batch runs in chunks of
 * floor(16 Mi / (2 hidden lt ly lx)) fields.  Kernels are enqueued on the
 * default stream: synchronise before reading y.  Calls that overlap on one
 * GPU share its scratch and race.  Nothing needs clearing between calls
 * (every scratch entry a call reads it has written).  Same arithmetic as
 * bafq_net_forward.  Returns 0, or: */
enum {
    BAFQ_NOALLOC_INVALID = 1,    /* null pointer, batch < 1 or invalid description */
    BAFQ_NOALLOC_TOO_LARGE = 2,  /* one field's activations exceed the scratch: 2 hidden lt ly lx > 16 Mi */
    BAFQ_NOALLOC_DEVICE = 3      /* a HIP error (device lookup or kernel launch) */
};
int bafq_net_forward_noalloc(const bafq_net_desc* d, const float* params, int batch, const float* x, float* y);
/* Device time in milliseconds of one forward pass over the first `batch`
 * inputs already on the device (the last forward or training input), the
 * average of `reps` passes between HIP events on the stream they run on. */
int bafq_net_time_forward(bafq_net* h, int batch, int reps, double* ms);
/* `steps` Adam steps (learning rate lr) on one batch; loss of the last step */
int bafq_net_train(bafq_net* h, int batch, const float* x, const float* target, int steps, float lr,
                   double* loss);

/* A learned global move of every chain on the device (no counterpart in the
 * reference; bafqmc_amd.global_moves.metropolis_global with
 * bafqmc_amd.proposal.LearnedProposal is the host-side form).  For a handle
 * with one model whose lattice is the network's (net may be NULL: no drift),
 * every chain's fields become phi' = phi + lam f(phi) + sigma[ch] xi, f the
 * network's output on phi (channels whose field does not couple are not
 * moved), xi standard normal numbers from a counter-based generator keyed by
 * seed, and G is rebuilt for phi' (as bafq_prepare).  out[c][0..4]:
 * log q(phi|phi') - log q(phi'|phi), sum phi^2, sum phi'^2, sum phi_2 and
 * sum phi'_2 (the Gaussian parts of log W).  bafq_global_finish then keeps
 * phi' where accept[c] != 0 and elsewhere restores phi, G, the checkpoint
 * lists and the counters from the device copy bafq_global_propose took. */
int bafq_global_propose(bafq_handle* h, bafq_net* net, double lam, const double* sigma, uint64_t seed,
                        double* out);
int bafq_global_finish(bafq_handle* h, const int32_t* accept);

/* A Hamiltonian global move of every chain with a learned force (no
 * counterpart in the reference).  With log W(phi) = -|phi - c|^2 / 2 + R(phi)
 * (c = center[ch] per channel), momenta p ~ N(0, 1) are drawn on the channels
 * whose field couples (counter-based generator keyed by seed) and `steps`
 * steps of split Hamiltonian dynamics are run: p += (angle/2) f(phi), (phi - c, p)
 * rotated by `angle`, p += (angle/2) f(phi'), with f = force_scale[ch] times
 * the network's output on phi (net may be NULL: f = 0, an exact draw from the
 * Gaussian part when steps * angle = pi/2).  G is rebuilt for phi' (as
 * bafq_prepare).
 * out[c][0..4]: (|p|^2 - |p'|^2) / 2, sum phi^2, sum phi'^2, sum phi_2 and
 * sum phi'_2 -- the layout of bafq_global_propose, so the exact acceptance is
 * the same formula; bafq_global_finish keeps or undoes the move.
 * 0 < angle <= 3.2, 1 <= steps <= 64. */
int bafq_hmc_propose(bafq_handle* h, bafq_net* net, double angle, int steps, const double* center,
                     const double* force_scale, uint64_t seed, double* out);

/* d log W / d phi for every field of every chain (laid out like the fields
 * of bafq_set_fields; W the weight of the global moves), from one left sweep
 * without updates: at each slice and channel, the diagonal of the Green's
 * function at the point of its local updates.  The chains' fields and RNG
 * streams are not changed.  No counterpart in the reference. */
int bafq_grad_log_weight(bafq_handle* h, double* grad);

#ifdef __cplusplus
}
#endif
#endif
