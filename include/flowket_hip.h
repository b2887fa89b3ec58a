/*
 * flowket_hip.h -- C ABI of the MI355X (gfx950) VMC hot-path library
 * (libflowket_hip.so).
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream passed
 * as `void*` (a hipStream_t; NULL = the null stream).  Work is enqueued on
 * that stream; no entry point synchronises unless stated.  Return value:
 * FK_OK (0) or an error code, with a message from fk_last_error().
 * Buffers are owned by the caller.  Calls that share a stream are ordered;
 * the library keeps no per-call state.
 *
 * The reference (FlowKet, Python/TensorFlow) has no C ABI: each entry point
 * below replaces the reference function cited next to it (paths relative to
 * FlowKet's src/flowket/), and the Python package flowket_amd binds them with
 * ctypes behind FlowKet's own class names.
 */
#ifndef FLOWKET_HIP_H
#define FLOWKET_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FK_OK 0
#define FK_ERR_ARG 1
#define FK_ERR_HIP 2
#define FK_ERR_UNSUPPORTED 3

/* Connection kinds of fk_find_conn. */
#define FK_CONN_SWAP 0          /* exchange spins a,b (Heisenberg)              */
#define FK_CONN_SWAP_ZERO_PAD 1 /* exchange a,b; unused slots all-zero (J1J2)   */
#define FK_CONN_FLIP 2          /* flip spin a (transverse-field Ising)         */

/* Arithmetic of the diagonal element in fk_find_conn. */
#define FK_DIAG_SUM_F64 0   /* sum_k diag_j[k]*s_a*s_b in double, table order   */
#define FK_DIAG_ISING_F32 1 /* Ising.find_conn's float32 rule (ising.py:25-41) */

/* Activations of ConvNetAutoregressive2D. */
#define FK_ACT_RELU 0
#define FK_ACT_ELU 1

/* Arithmetic of the fused network kernels: FP32 MFMA, or 16-bit MFMA
 * (bf16 / fp16 weights and activations) with FP32 accumulation.  fp16 runs at
 * the bf16 rate with 8x finer rounding; its range is |x| < 65504. */
#define FK_PREC_FP32 0
#define FK_PREC_BF16 1
#define FK_PREC_FP16 2

int fk_version(void);
const char* fk_last_error(void);

/* ---------------------------------------------------------------------------
 * Operators: connected configurations and matrix elements.
 * Replaces HeisenbergFindConn.find_conn (operators/heisenberg.py:72-121),
 * Ising.find_conn (operators/ising.py:17-46) and the J1J2 netket wrapper
 * (operators/j1j2.py:7-69, operators/netket_operator.py:46-66).
 *
 * The operator is described by a connection table built once on the host:
 *   conn_a[c], conn_b[c]  (c = 0..n_conn-1; entry 0 is the diagonal, ignored)
 *       sites of connection c, -1 for "no bond" (use=0, mel=0)
 *   conn_mel[c]           off-diagonal element of connection c when it applies
 *   diag_a/diag_b/diag_j  diagonal bond table; with diag_rule
 *       FK_DIAG_SUM_F64    sum_k diag_j[k]*s[diag_a[k]]*s[diag_b[k]]
 *                          accumulated in double in k order (Heisenberg,
 *                          J1J2: exact for their integer/dyadic couplings);
 *       FK_DIAG_ISING_F32  the reference's numpy float32 arithmetic
 *                          (operators/ising.py:25,30-41): the entries are
 *                          grouped by site (diag_a non-decreasing), each
 *                          site's value starts at 0.0f and takes
 *                          v = (float)((double)v + diag_j[k]*s_a*s_b) per
 *                          entry in table order, and the n_sites values are
 *                          summed in float32 in numpy's pairwise order
 *                          (the ndarray.sum over the lattice axes, :41).
 * spins: (batch, n_sites) int8 +-1.
 * Outputs (any may be NULL except mel/use):
 *   conn_out (n_conn, batch, n_sites) int8, mel_out (n_conn, batch) double,
 *   use_out (n_conn, batch) uint8, n_used_out (batch) int32.
 * ------------------------------------------------------------------------- */
int fk_find_conn(const int8_t* spins, int64_t batch, int32_t n_sites,
                 int32_t kind, const int32_t* conn_a, const int32_t* conn_b,
                 const double* conn_mel, int32_t n_conn,
                 const int32_t* diag_a, const int32_t* diag_b,
                 const double* diag_j, int32_t n_diag, int32_t diag_rule,
                 int8_t* conn_out, double* mel_out, uint8_t* use_out,
                 int32_t* n_used_out, void* stream);

/* Ragged index map of the unbalanced local-energy path
 * (observables/monte_carlo/operator.py:6-10,26-30): from use (n_conn, batch)
 * writes offsets (batch+1, exclusive prefix sum of used counts, offsets[batch]
 * = total) and pairs (total, 2) = (sample, connection) in sample-major,
 * connection-minor order.  `pairs` must hold n_conn*batch entries.
 * `scratch` must hold fk_compact_scratch_bytes(batch) bytes. */
int64_t fk_compact_scratch_bytes(int64_t batch);
int fk_compact_connections(const uint8_t* use, int64_t batch, int32_t n_conn,
                           int64_t* offsets, int32_t* pairs, void* scratch,
                           void* stream);

/* The same pairs regrouped by a per-slot bucket (bucket_of[c] in
 * [0, n_buckets), n_buckets <= 128; the row-reuse forward buckets a
 * connection by the first lattice row it changes): pairs (total, 2) in
 * (bucket, sample, slot) order, out_index[k] = the position of pair k in the
 * sample-major order of fk_compact_connections (offsets from it),
 * bucket_offsets (n_buckets + 1) the start of every bucket and the total.
 * `scratch` holds fk_bucket_scratch_bytes(batch, n_buckets) bytes.  No
 * counterpart in the reference (its E_loc evaluates every connection with a
 * full forward). */
int64_t fk_bucket_scratch_bytes(int64_t batch, int32_t n_buckets);
int fk_bucket_connections(const uint8_t* use, int64_t batch, int32_t n_conn,
                          const int32_t* bucket_of, int32_t n_buckets,
                          const int64_t* offsets, int32_t* pairs,
                          int64_t* out_index, int64_t* bucket_offsets,
                          void* scratch, void* stream);

/* E_loc (observables/monte_carlo/operator.py:13-33): for sample b,
 *   sum_k mel[c_k, b] * exp(logpsi[k] - logpsi[offsets[b]])
 * over k in [offsets[b], offsets[b+1]), difference and exp in complex64 as the
 * reference, product and sum in complex128.  logpsi (total, 2) float (re, im);
 * eloc_out (batch, 2) double.  mel (n_conn, batch) holds the real parts of
 * the matrix elements; mel_im, when not NULL, their imaginary parts (the
 * complex128 elements of operators/netket_operator.py:56-60). */
int fk_local_energy(const int64_t* offsets, const int32_t* pairs,
                    const float* logpsi, const double* mel,
                    const double* mel_im, int64_t batch, double* eloc_out,
                    void* stream);

/* Balanced path (operator.py:35-42): logpsi (n_conn, batch, 2) for every
 * slot, eloc_out (batch, 2); mel / mel_im as for fk_local_energy. */
int fk_local_energy_dense(const float* logpsi, const double* mel,
                          const double* mel_im, int32_t n_conn, int64_t batch,
                          double* eloc_out, void* stream);

/* BaseObservable.estimate (observables/monte_carlo/observable.py:10-14):
 * out[0..3] = (Re mean, Im mean, population variance of Re, batch) in double,
 * reduced on the device in two passes. */
int fk_energy_moments(const double* eloc, int64_t batch, double* out,
                      void* stream);

/* SigmaZ / AbsSigmaZ local values (observables/monte_carlo/sigma_z.py:7-20):
 * out[2b] = (sum of the n_sites spins of sample b) / n_sites, or its absolute
 * value when `absolute` is non-zero; out[2b+1] = 0 (complex pairs, as
 * fk_energy_moments reads them). */
int fk_sigma_z(const int8_t* spins, int64_t batch, int32_t n_sites,
               int32_t absolute, double* out, void* stream);

/* ---------------------------------------------------------------------------
 * ConvNetAutoregressive2D (machines/conv_net_autoregressive_2D.py:8-78 with
 * AutoNormalizedAutoregressiveMachine, machines/abstract_machine.py:49-71).
 *
 * Parameters live in a flat float32 vector in network order (per causal
 * block: vconv k,b; hconv k,b; x1x1 k,b; y1x1 k,b; h3conv k,b; then head k,b;
 * Keras kernel layout (kh, kw, cin, cout)).  The kernels read them from two
 * packed device buffers (MFMA fragment order + an fp32 side buffer) produced
 * by fk_pack_gather with the maps from fk_convar_pack_maps.
 *
 * Channels: any even count up to 64 (bf16) / 2..64 (fp32).  The kernels are
 * compiled for 32 and 64 (bf16) and 16, 32, 64 (fp32) channels; other widths
 * run on the next wider kernel with zero weights and biases in the padding
 * channels (exactly the narrower network's result, at the wider cost).
 * ------------------------------------------------------------------------- */

/* Host-only.  Sizes (elements) of the packed buffers; returns FK_OK or
 * FK_ERR_UNSUPPORTED for shapes the kernels do not implement. */
int fk_convar_packed_sizes(int32_t depth, int32_t channels, int32_t precision,
                           int64_t* n_weights, int64_t* n_aux);
/* Host-only.  Fills wmap (n_weights) and amap (n_aux) with source indices
 * into the flat parameter vector (-1 = zero). */
int fk_convar_pack_maps(int32_t depth, int32_t channels, int32_t precision,
                        int32_t* wmap, int32_t* amap);
/* Gathers params[map[i]] (0 where map[i] < 0) into dst as float32
 * (precision FP32) or bfloat16 round-to-nearest-even (BF16). */
int fk_pack_gather(const float* params, const int32_t* map, int64_t n,
                   int32_t precision, void* dst, void* stream);

/* Fused forward of the whole network.  Configurations are either the rows of
 * `spins` (pairs == NULL: m_count rows) or connected configurations
 * (pairs (m, 2) = (sample, connection) applied to spins with the connection
 * table conn_a/conn_b of `conn_kind`).  If m_dev != NULL the count is read
 * from device memory (*m_dev, int64), m_count is then an upper bound used to
 * size the launch.  Outputs: logpsi (m, 2) float (re, im); cond_logp
 * (m, H, W, 2) float = 2*Re of the normalised conditionals, or NULL. */
int fk_convar_forward(const void* wpacked, const float* aux, int32_t precision,
                      int32_t H, int32_t W, int32_t depth, int32_t channels,
                      int32_t activation, const int8_t* spins,
                      const int32_t* pairs, const int64_t* m_dev,
                      int64_t m_count, int32_t conn_kind,
                      const int32_t* conn_a, const int32_t* conn_b,
                      float* logpsi, float* cond_logp, void* stream);

/* Same as fk_convar_forward, for every lattice the kernels support: when the
 * five activation lattices of one configuration exceed the LDS (e.g. 20x20
 * with 64 bf16 channels), the horizontal-stack lattices are kept in
 * `workspace`, a caller-owned device buffer of at least
 * fk_convar_forward_workspace_bytes(H, W, depth, channels, activation,
 * precision) bytes (0 when
 * everything fits in LDS; then workspace may be NULL).  Calls that share a
 * workspace must be ordered on one stream.  fk_convar_forward returns
 * FK_ERR_ARG for such lattices.  (No counterpart in the reference: its
 * Keras convolutions have no on-chip memory limit.) */
int fk_convar_forward_ws(const void* wpacked, const float* aux, int32_t precision,
                         int32_t H, int32_t W, int32_t depth, int32_t channels,
                         int32_t activation, const int8_t* spins,
                         const int32_t* pairs, const int64_t* m_dev,
                         int64_t m_count, int32_t conn_kind,
                         const int32_t* conn_a, const int32_t* conn_b,
                         float* logpsi, float* cond_logp, void* workspace,
                         int64_t workspace_bytes, void* stream);
/* Row reuse for connected configurations (no counterpart in the reference,
 * which runs a full forward per connection; the results are the same up to
 * FP32 summation order).  Every activation of lattice row r depends on input
 * rows <= r only, so a configuration equal to its sample on the rows before
 * row0 shares the sample's activations there:
 *   sample pass (row0 = 0, cache_write = 1, pairs = NULL): the forward of
 *     m_count rows of `spins`, storing for every sample its blocks' vertical
 *     inputs and concatenations in rcache (fk_convar_rcache_bytes(...) bytes
 *     per sample) and its per-row prefix sums of the site log-probabilities
 *     in prefix ((batch, H + 1, 2) floats);
 *   row pass (2 <= row0 < H, pairs given): configurations whose rows before
 *     row0 equal their sample's; only rows row0 .. H-1 are computed, the two
 *     rows above them are copied from the sample's cache and the prefix sum
 *     of the rows before row0 is added;
 *   row0 = 0, cache_write = 0: a plain pass.
 * m_range (may be NULL) is a device pointer to (first, end): the launch
 * evaluates pairs[first .. end) (m_count bounds end - first and sizes the
 * grid), so buckets of fk_bucket_connections need no host round trip.
 * out_index (may be NULL) places the result of configuration m at row
 * out_index[first + m] of logpsi.  3 <= H <= 64.  Workspace as fk_convar_forward_ws
 * (its size covers the row passes). */
int64_t fk_convar_rcache_bytes(int32_t H, int32_t W, int32_t depth,
                               int32_t channels, int32_t precision);
int fk_convar_forward_rows(const void* wpacked, const float* aux,
                           int32_t precision, int32_t H, int32_t W,
                           int32_t depth, int32_t channels, int32_t activation,
                           const int8_t* spins, const int32_t* pairs,
                           const int64_t* m_range, int64_t m_count,
                           int32_t conn_kind, const int32_t* conn_a,
                           const int32_t* conn_b, int32_t row0, void* rcache,
                           int32_t cache_write, float* prefix,
                           const int64_t* out_index, float* logpsi,
                           void* workspace, int64_t workspace_bytes,
                           void* stream);

/* Same as fk_convar_forward_ws with no workspace argument and no allocation
 * of any kind.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The large-lattice scratch is a
 * static device buffer of fk_convar_forward_noalloc_scratch_bytes() bytes
 * (32 MiB) per GPU, held until the process ends; a network whose
 * fk_convar_forward_workspace_bytes exceeds it is refused with FK_ERR_ARG.
 * Every call on one GPU shares that buffer: calls that overlap on one GPU
 * (different streams, or host threads) race and must be ordered by the
 * caller.  The kernels initialise what they read from it, so no clearing is
 * needed between calls. */
int fk_convar_forward_noalloc(const void* wpacked, const float* aux,
                              int32_t precision, int32_t H, int32_t W,
                              int32_t depth, int32_t channels,
                              int32_t activation, const int8_t* spins,
                              const int32_t* pairs, const int64_t* m_dev,
                              int64_t m_count, int32_t conn_kind,
                              const int32_t* conn_a, const int32_t* conn_b,
                              float* logpsi, float* cond_logp, void* stream);
int64_t fk_convar_forward_noalloc_scratch_bytes(void);

/* Host-only (queries the current device's CU count).  Bytes of workspace
 * fk_convar_forward_ws needs for a network (under its selected schedule),
 * 0 if none, -1 if the lattice is unsupported. */
int64_t fk_convar_forward_workspace_bytes(int32_t H, int32_t W, int32_t depth,
                                          int32_t channels, int32_t activation,
                                          int32_t precision);

/* Launch schedules of fk_convar_forward (no counterpart in the reference:
 * the knobs of this library's autotuner).  Each (channels, precision) has
 * fk_convar_num_schedules() compiled schedules (loop order tile_outer 0/1,
 * waves per workgroup); fk_convar_set_schedule selects one for a network
 * (lattice H x W, depth, channels, activation, precision; -1 restores the
 * default, index 0) for every later forward launch of that network in the
 * process; results are identical up to FP32 summation order.
 * Host-only, thread-safe. */
int32_t fk_convar_num_schedules(int32_t channels, int32_t precision);
int fk_convar_schedule_info(int32_t channels, int32_t precision,
                            int32_t schedule, int32_t* tile_outer,
                            int32_t* waves);
int fk_convar_set_schedule(int32_t H, int32_t W, int32_t depth,
                           int32_t channels, int32_t activation,
                           int32_t precision, int32_t schedule);
int32_t fk_convar_get_schedule(int32_t H, int32_t W, int32_t depth,
                               int32_t channels, int32_t activation,
                               int32_t precision);

/* Plain autoregressive draw at one site (deepar/samplers/autoregressive.py:
 * 42-44): spins[b, site] = +1 if exp(cond_logp[b, site, 0]) > u[b, site]
 * else -1, for all b; cond_logp (batch, n_sites, 2) float, u (batch, n_sites)
 * double, spins (batch, n_sites) int8. */
int fk_ar_draw_site(const float* cond_logp, const double* uniforms,
                    int64_t batch, int32_t n_sites, int32_t site,
                    int8_t* spins, void* stream);

/* ---------------------------------------------------------------------------
 * Fast autoregressive sampling (deepar/samplers/fast_autoregressive.py:13-76):
 * each (layer, site) activation computed once, spin drawn at every site with
 * the AutoregressiveSampler rule s = +1 iff exp(log p(+1)) > u
 * (deepar/samplers/autoregressive.py:42-44) from injected uniforms
 * (batch, H*W) double.  Output spins (batch, H*W) int8; cond_logp (batch,
 * H*W, 2) float = log-probabilities of each site's conditional, or NULL.
 * `workspace` holds fk_fast_sampler_workspace_bytes(...) bytes (the
 * activation cache; contents need no initialisation).  Channels 17..64
 * (bf16: 2..64), run on the 32- or 64-channel kernel as above.
 * ------------------------------------------------------------------------- */
int64_t fk_fast_sampler_workspace_bytes(int32_t H, int32_t W, int32_t depth,
                                        int32_t channels, int32_t precision,
                                        int64_t batch);
int fk_fast_sample(const void* wpacked, const float* aux, int32_t precision,
                   int32_t H, int32_t W, int32_t depth, int32_t channels,
                   int32_t activation, int64_t batch, const double* uniforms,
                   void* workspace, int8_t* out, float* cond_logp, void* stream);

/* ---------------------------------------------------------------------------
 * Per-sample log-psi gradients O = d log psi / d theta (real parameters,
 * complex log psi): Machine.predictions_jacobian (machines/abstract_machine.py
 * :24-28) re-paired as complex O_k (optimizers/complex_values_optimizer.py:
 * 8-9), per example as utils/jacobian.py:114-123.  params: flat FP32
 * parameters (effective kernels); o_re / o_im (batch, row_stride) FP32, the
 * first P entries of each row in the flat order (row_stride 0 means P);
 * logpsi (batch, 2) or NULL.  FP32 throughout.  `workspace` holds
 * fk_convar_grad_workspace_bytes(...) bytes.
 * ------------------------------------------------------------------------- */
int64_t fk_convar_grad_workspace_bytes(int32_t H, int32_t W, int32_t depth,
                                       int32_t channels, int64_t batch);
int fk_convar_grad(const float* params, int32_t H, int32_t W, int32_t depth,
                   int32_t channels, int32_t activation, const int8_t* spins,
                   int64_t batch, void* workspace, float* o_re, float* o_im,
                   int64_t row_stride, float* logpsi, void* stream);

/* Implicit products with O, never forming it (the
 * compute_jvp_instead_of_full_jacobian path of optimizers/
 * stochastic_reconfiguration/optimizer.py:116-124 through
 * optimizers/utils.py:5-34, and the batch gradient of the VMC loss,
 * optimization/loss.py:4-5).  Channels a multiple of 32; FP32 MFMA.
 *   fk_convar_vjp: grad = scale * sum_b (cot[b,0] d Re log psi_b / d theta
 *                  + cot[b,1] d Im log psi_b / d theta), P floats, i.e.
 *                  scale * Re[O^H u] for u = cot[:,0] + i cot[:,1].  One
 *                  backward pass per sample; the samples' contributions are
 *                  summed in per-workgroup partial rows and reduced in a
 *                  fixed order in FP64 (deterministic).
 *   fk_convar_jvp: jvp (batch, 2) = (d Re log psi_b . t, d Im log psi_b . t)
 *                  for a parameter direction t (P floats), i.e. O t, by
 *                  forward-mode differentiation of the network.
 * logpsi (batch, 2) or NULL.  Workspaces hold *_workspace_bytes(...) bytes
 * (host-only queries: they depend on the device's CU count). */
int64_t fk_convar_vjp_workspace_bytes(int32_t H, int32_t W, int32_t depth,
                                      int32_t channels, int64_t batch);
int fk_convar_vjp(const float* params, int32_t H, int32_t W, int32_t depth,
                  int32_t channels, int32_t activation, const int8_t* spins,
                  int64_t batch, const float* cot, double scale,
                  void* workspace, float* grad, float* logpsi, void* stream);
int64_t fk_convar_jvp_workspace_bytes(int32_t H, int32_t W, int32_t depth,
                                      int32_t channels, int64_t batch);
int fk_convar_jvp(const float* params, const float* tangent, int32_t H,
                  int32_t W, int32_t depth, int32_t channels,
                  int32_t activation, const int8_t* spins, int64_t batch,
                  void* workspace, float* jvp, float* logpsi, void* stream);

/* Weight normalisation chain rule (deepar/layers/wrappers.py:123-135):
 * a wrapped layer's kernel is k = v/n*s per output channel, n =
 * sqrt(max(sum v^2, 1e-12)) (tf.nn.l2_normalize), s = exp(g) when exp_norm
 * else g.  `layers` holds 4 int64 per wrapped layer: kernel offset in the
 * flat parameters, kh*kw*cin, cout (<= max_cout), offset of its g.  params
 * are the flat parameters with v in the kernel slots, g the g vector.
 *   fk_wn_grad_rows: in place on rows (nrows, ld >= P + n_g): each row's
 *                    d/dk in [0, P) becomes d/dv in the kernel slots, and
 *                    d/dg is written at P + goff + c (the trainable order
 *                    [params | g]).
 *   fk_wn_tangent:   t_eff (P) = dk for a direction t over [params | g]. */
int fk_wn_grad_rows(float* rows, int64_t nrows, int64_t ld,
                    const float* params, const float* g, const int64_t* layers,
                    int32_t n_layers, int32_t max_cout, int32_t exp_norm,
                    int64_t n_params, void* stream);
int fk_wn_tangent(const float* params, const float* g, const int64_t* layers,
                  int32_t n_layers, int32_t max_cout, int32_t exp_norm,
                  const float* t, float* t_eff, int64_t n_params,
                  void* stream);

/* ---------------------------------------------------------------------------
 * Stochastic reconfiguration on the implicit S = Re[(O-Obar)^H (O-Obar)]/B
 * + shift*I (optimizers/stochastic_reconfiguration/optimizer.py:69-92, force
 * :94-109; CG of linear_equations.py:34-137 is driven by the host with these
 * pieces).  O as fp32 (re, im) arrays (batch, n_params); products FP32,
 * reductions FP64.
 *   fk_o_mean:     Obar = column means.
 *   fk_o_times_v:  u (batch, 2) = (O - Obar) v, v real (n_params).
 *   fk_oh_times_u: out = scale * Re[(O - Obar)^H u] + shift * v (v may be
 *                  NULL: no shift term); scratch2 holds 2 doubles.
 *   fk_dot_f64:    *out = sum a_i b_i in FP64; scratch holds
 *                  fk_sr_scratch_doubles() doubles (also the scratch of
 *                  fk_o_times_v).
 *   fk_axpy:       y += alpha x.   fk_xpby: p = r + beta p.
 * ------------------------------------------------------------------------- */
int fk_o_mean(const float* o_re, const float* o_im, int64_t batch,
              int64_t n_params, float* mean_re, float* mean_im, void* stream);
int fk_o_times_v(const float* o_re, const float* o_im, const float* mean_re,
                 const float* mean_im, const float* v, int64_t batch,
                 int64_t n_params, float* u, double* scratch, void* stream);
int fk_oh_times_u(const float* o_re, const float* o_im, const float* mean_re,
                  const float* mean_im, const float* u, int64_t batch,
                  int64_t n_params, double scale, double shift, const float* v,
                  float* out, double* scratch2, void* stream);
int64_t fk_sr_scratch_doubles(void);
/* out = scale * Re[(O - Obar)^H (O - Obar) v] + shift * v, the S v of
 * optimizer.py:78-81 on a stored O.  block_rows = 0: fk_o_times_v then
 * fk_oh_times_u (two passes over O, the faster form on MI355X).  block_rows
 * in [1, 32]: one pass over O from HBM -- the rows go in blocks of
 * block_rows; each launch forms the partial dots O v of one block and adds
 * Re[O^H u] of the previous one, whose rows are still in the last-level
 * cache.  u (batch, 2) receives (O - Obar) v as fk_o_times_v's.  workspace:
 * fk_sr_matvec_workspace_bytes(batch, n_params) bytes, 16-byte aligned (may
 * be NULL for block_rows = 0); scratch: fk_sr_scratch_doubles() doubles.
 * n_params not a multiple of 4 always takes the two-pass form. */
int64_t fk_sr_matvec_workspace_bytes(int64_t batch, int64_t n_params);
int fk_sr_matvec(const float* o_re, const float* o_im, const float* mean_re,
                 const float* mean_im, const float* v, int64_t batch,
                 int64_t n_params, double scale, double shift, float* out,
                 float* u, int32_t block_rows, void* workspace,
                 int64_t workspace_bytes, double* scratch, void* stream);
int fk_dot_f64(const float* a, const float* b, int64_t n, double* scratch,
               double* out, void* stream);
int fk_axpy(double alpha, const float* x, float* y, int64_t n, void* stream);

/* Conjugate gradient of linear_equations.py:34-137 (no preconditioner) with
 * its scalars on the device, so a solve needs no host round trip (and can be
 * captured in a graph): fk_cg_init sets x = 0, r = p = rhs and the state
 * (fk_cg_state_doubles() doubles: r.r, p.z, new r.r, tol*||rhs|| with both
 * factors rounded to float32, done flag, iterations, max_iter); after each
 * matrix product z = A p the caller launches fk_cg_step, which performs one
 * iteration unless the reference's loop would already have stopped
 * (i >= max_iter or ||r|| <= tol*||rhs||), leaving everything unchanged then.
 * state[5] is the iteration count, sqrt(state[0]) the residual norm.
 * Dots in FP64 (scratch as fk_dot_f64). */
int64_t fk_cg_state_doubles(void);
int fk_cg_init(const float* rhs, float* x, float* r, float* p, int64_t n,
               double tol, int64_t max_iter, double* state, double* scratch,
               void* stream);
int fk_cg_step(float* x, float* r, float* p, const float* z, int64_t n,
               double* state, double* scratch, void* stream);

/* u (batch, 2) complex pairs: u -= mean(u) (FP64 sums; scratch holds 2
 * doubles).  Centres O t for the implicit S v = Re[(O-Obar)^H (O-Obar) v]/B. */
int fk_center_pairs(float* u, int64_t batch, double* scratch, void* stream);
int fk_xpby(const float* r, double beta, float* p, int64_t n, void* stream);

/* Adam step t (>= 1) of the Keras/TensorFlow optimizer the reference's
 * examples train with (tf.keras.optimizers.Adam, examples/
 * basic_autoregressive_2d.py:24): lr_t = lr*sqrt(1-beta_2^t)/(1-beta_1^t),
 * m += (g-m)(1-beta_1), v += (g^2-v)(1-beta_2),
 * params -= lr_t*m/(sqrt(v)+epsilon), elementwise in FP32 on n values. */
int fk_adam_step(float* params, float* m, float* v, const float* grad,
                 int64_t n, double lr, double beta_1, double beta_2,
                 double epsilon, int64_t t, void* stream);

/* ---------------------------------------------------------------------------
 * Complex RBM (machines/rbm.py:35-48): log psi = sum_j a_j s_j +
 * sum_h lncosh(b_h + sum_j W_jh s_j) with each complex parameter stored as
 * (real, imag) and used as real - 1j*imag (layers/complex/base_layer.py:
 * 30-35).  W (n_visible, n_hidden) row-major.  Configurations as in
 * fk_convar_forward (rows of spins, or (sample, connection) pairs with a
 * connection table).  logpsi (m, 2) float.
 * ------------------------------------------------------------------------- */
int fk_rbm_log_psi(const float* w_re, const float* w_im, const float* b_re,
                   const float* b_im, const float* a_re, const float* a_im,
                   int32_t n_visible, int32_t n_hidden, const int8_t* spins,
                   const int32_t* pairs, const int64_t* m_dev, int64_t m_count,
                   int32_t conn_kind, const int32_t* conn_a,
                   const int32_t* conn_b, float* logpsi, void* stream);

/* Per-sample O = d log psi / d(parameter) of the RBM over its stored real
 * and imaginary parts (the Jacobian of complex layers that
 * optimizers/complex_values_optimizer.py:60-76 takes): with T = tanh(theta),
 * d/dW_re[j,h] = s_j T_h, d/dW_im = -i s_j T_h, d/db_re = T_h, d/db_im =
 * -i T_h, d/da_re = s_j, d/da_im = -i s_j.  o_re / o_im (batch, row_stride)
 * FP32 in the flat order W_re, W_im, b_re, b_im, a_re, a_im (row_stride 0 =
 * the parameter count). */
int fk_rbm_grad(const float* w_re, const float* w_im, const float* b_re,
                const float* b_im, const float* a_re, const float* a_im,
                int32_t n_visible, int32_t n_hidden, const int8_t* spins,
                int64_t batch, float* o_re, float* o_im, int64_t row_stride,
                void* stream);

/* Metropolis-Hastings chains of an RBM (samplers/metropolis_hastings.py:
 * 15-145): kind 0 = MetropolisHastingsLocal (flip one site), 1 =
 * MetropolisHastingsExchange (swap a site with a neighbour offset in
 * {-1,0,1}^2, periodic).  Each chain runs n_burn sweeps, then records
 * n_record samples every (unused + 1) sweeps into out (chains, n_record, N);
 * one sweep = one proposal per chain, accepted iff
 * exp(min(2 Re(logpsi' - logpsi), 0)) > U.  state (chains, N) is read and
 * updated; accepts (chains) counts accepted proposals of the recorded sweeps;
 * logpsi (chains, 2) final values or NULL.  Randomness: injected
 * (steps, chains) arrays r_site (flat site), r_move ((dy+1)*3+(dx+1),
 * exchange only) and r_u (uniforms), or, when r_u is NULL, a counter-based
 * generator keyed by `seed`.  A proposal whose log psi is not finite is
 * rejected and sets *nonfinite = 1 (nonfinite may be NULL; the reference
 * raises, samplers/metropolis_hastings.py:103-105); the same holds for
 * fk_mh_accept and fk_mh_accept_independent. */
int fk_mh_rbm(const float* w_re, const float* w_im, const float* b_re,
              const float* b_im, const float* a_re, const float* a_im,
              int32_t n_visible, int32_t n_hidden, int32_t H, int32_t W,
              int32_t kind, int64_t chains, int64_t n_burn, int64_t n_record,
              int64_t unused, int8_t* state, int8_t* out, int64_t* accepts,
              float* logpsi, const int32_t* r_site, const int32_t* r_move,
              const double* r_u, uint64_t seed, int32_t* nonfinite,
              void* stream);

/* Machine-agnostic MH steps (MetropolisHastingsSymmetricProposal._sweep,
 * samplers/metropolis_hastings.py:99-113): candidates from the chain states
 * (kind / r_site / r_move as fk_mh_rbm, one entry per chain), then
 * acceptance given the candidates' log psi; accepted (chains) int32. */
int fk_mh_propose(const int8_t* state, int64_t chains, int32_t H, int32_t W,
                  int32_t kind, const int32_t* r_site, const int32_t* r_move,
                  int8_t* candidates, void* stream);
int fk_mh_accept(int8_t* state, float* logpsi_cur, const int8_t* candidates,
                 const float* logpsi_cand, const double* uniforms,
                 int64_t chains, int32_t n_sites, int32_t* accepted,
                 int32_t* nonfinite, void* stream);

/* Independent-proposal acceptance with the exact ratio (no reference
 * counterpart: MetropolisHastingsGlobal, samplers/metropolis_hastings.py:
 * 148-157, treats any proposal as symmetric): accept iff
 * exp(min(2 Re(logpsi' - logpsi) + log q(x) - log q(x'), 0)) > U. */
int fk_mh_accept_independent(int8_t* state, float* logpsi_cur, float* logq_cur,
                             const int8_t* candidates, const float* logpsi_cand,
                             const float* logq_cand, const double* uniforms,
                             int64_t chains, int32_t n_sites, int32_t* accepted,
                             int32_t* nonfinite, void* stream);

/* ---------------------------------------------------------------------------
 * Learned autoregressive proposal (no reference counterpart): a masked MLP
 * (MADE) n_sites -> hidden -> hidden -> n_sites logits, ReLU, BF16 weights
 * (w1 (hidden, n_sites), w2t (hidden, hidden) = W2 transposed,
 * w3 (n_sites, hidden)), FP32 biases and accumulation, hidden activations
 * rounded to BF16.  Site i is +1 with probability sigmoid(logit_i) given
 * sites < i.
 *   fk_made_pack:     writes the degree-sorted, masked weights that log q and
 *                     sampling read (fk_made_packed_bytes bytes); call after
 *                     every change of the weights.
 *   fk_made_log_prob: logq (batch) = log q(x).
 *   fk_made_sample:   x ~ q from uniforms (batch, n_sites), with log q(x)
 *                     (the same arithmetic as fk_made_log_prob).
 *   fk_made_fit_step: one full-batch Adam step of -mean log q(x) on the FP32
 *                     master vector [w1 | b1 | w2t | b2 | w3 | b3] (b1, b2,
 *                     b3 point into it), refreshing the BF16 copies;
 *                     workspace holds fk_made_fit_workspace_floats floats.
 * ------------------------------------------------------------------------- */
int64_t fk_made_packed_bytes(int32_t n_sites, int32_t hidden);
int fk_made_pack(const void* w1, const float* b1, const void* w2t,
                 const float* b2, const void* w3, const float* b3,
                 int32_t n_sites, int32_t hidden, void* packed, void* stream);
int fk_made_log_prob(const void* packed, int32_t n_sites, int32_t hidden,
                     const int8_t* x, int64_t batch, float* logq, void* stream);
int fk_made_sample(const void* packed, int32_t n_sites, int32_t hidden,
                   const double* uniforms, int64_t batch, int8_t* x,
                   float* logq, void* stream);
int64_t fk_made_fit_workspace_floats(int32_t n_sites, int32_t hidden,
                                     int64_t batch);
int fk_made_fit_step(float* master, float* adam_m, float* adam_v, void* w1,
                     float* b1, void* w2t, float* b2, void* w3, float* b3,
                     int32_t n_sites, int32_t hidden, const int8_t* x,
                     int64_t batch, float* workspace, float lr, float beta1,
                     float beta2, float eps, int64_t step, void* stream);

/* ---------------------------------------------------------------------------
 * Exact enumeration over all 2^n_sites configurations (state index k: site i
 * is +1 when bit i of k is set, else -1 -- exact/utils.py:49-58).
 *   fk_exact_states:        the configurations first .. first+count-1,
 *                           int8 (count, n_sites).  Replaces the `states`
 *                           table of ExactVariational._build_wave_function_arrays
 *                           (optimization/exact_variational.py:104-117).
 *   fk_exact_normalize:     probs[k] = |psi_k|^2 / sum |psi|^2 from log psi
 *                           (n_states, 2) float32; scalars[0] = log sum
 *                           |psi|^2, scalars[1] = max 2 Re log psi.
 *                           ExactVariational._update_wave_function_arrays
 *                           (exact_variational.py:127-133) with
 *                           complex_norm_log_fsum_exp (exact/utils.py:79-82).
 *   fk_exact_local_energy:  eloc (n_states, 2) float64 = sum over used slots
 *                           of mel * psi(k') / psi(k), from fk_find_conn's mel
 *                           and use over the enumerated states, the connected
 *                           configurations located by index (no network
 *                           evaluation).  ExactObservable.update_local_energy
 *                           (exact_variational.py:80-95, its naive energies).
 *   fk_exact_energy:        moments[0..1] = E = sum p E_loc, moments[2] =
 *                           sum p (Re E_loc - Re E)^2; coef (n_states, 2) =
 *                           p conj(E_loc - E) when coef is not NULL
 *                           (exact_variational.py:86-100,135-138).
 *   fk_exact_sample:        batch configurations drawn from probs by the
 *                           inverse CDF of the uniforms (batch) float64 --
 *                           ExactSampler.__next__ (samplers/exact_sampler.py:14-20).
 * Workspace: fk_exact_workspace_doubles(n_states) doubles, shared by
 * normalize, energy and sample (which uses the n_states-long tail for the CDF).
 * ------------------------------------------------------------------------- */
int fk_exact_states(int32_t n_sites, int64_t first, int64_t count,
                    int8_t* states, void* stream);
int64_t fk_exact_workspace_doubles(int64_t n_states);
int fk_exact_normalize(const float* logpsi, int64_t n_states, double* probs,
                       double* scalars, double* workspace, void* stream);
int fk_exact_local_energy(const float* logpsi, int64_t n_states,
                          int32_t n_sites, int32_t kind, const int32_t* conn_a,
                          const int32_t* conn_b, const double* mel,
                          const uint8_t* use, int32_t n_conn, double* eloc,
                          void* stream);
int fk_exact_energy(const double* probs, const double* eloc, int64_t n_states,
                    double* coef, double* moments, double* workspace,
                    void* stream);
int fk_exact_sample(const double* probs, int64_t n_states, int32_t n_sites,
                    const double* uniforms, int64_t batch, int8_t* out,
                    double* workspace, void* stream);

/* Symmetrised wave functions (machines/ensemble.py:14-69).
 *   fk_transform_spins:  out[b, q] = sign * in[b, src[q]] for a batch of
 *                        (batch, n_sites) int8 configurations (the lattice
 *                        symmetries Rot90 / FlipLeftRight of
 *                        layers/dihedral_4_invariants.py:8-30 and the spin
 *                        flip x -> -x of ensemble.py:65-69 as a gather).
 *   fk_ensemble_combine: the probabilistic ensemble of K log psi blocks of
 *                        `stride` (re, im) float32 rows (ensemble.py:14-20):
 *                        Re = 0.5 logsumexp(2 Re_k) - 0.5 log K,
 *                        Im = arg(mean exp(i Im_k)); computed in FP64.  Only
 *                        the first min(*m_dev, m) rows are written when m_dev
 *                        is non-NULL. */
int fk_transform_spins(const int8_t* in, int64_t batch, int32_t n_sites,
                       const int32_t* src, int32_t sign, int8_t* out,
                       void* stream);
int fk_ensemble_combine(const float* logpsi, int32_t k_members, int64_t stride,
                        int64_t m, const int64_t* m_dev, float* out,
                        void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLOWKET_HIP_H */
