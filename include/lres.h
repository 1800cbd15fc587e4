/* C-ABI of the MI355X LRes window step (liblres_hip.so).
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success and a
 * non-zero status otherwise; lres_last_error() then gives the message (the
 * Python host layer raises RuntimeError with it, the exception type the
 * reference's pybind11 module raises for LAPACK failures, src/lrft_cpp.cpp:94,110).
 * Nothing throws across this boundary.
 *
 * Pointer conventions: arguments named h_* are host memory, d_* device memory
 * (caller-owned, e.g. torch tensors).  All device work is ordered on the
 * stream passed to lres_plan_create (0 = the null stream); entry points that
 * return host scalars synchronise that stream.
 *
 * Reference interface replaced: the pybind11 module lrft._lrft_cpp
 * (src/lrft_cpp.cpp:399-404), function lrft_windows (src/lrft_cpp.cpp:129-233),
 * and the per-geometry set-up that LRFTLCAO/LRFTLCAOCpp do in Python
 * (lrft/lcao3d.py:132-161, 343-357).  The eigen route of the reference becomes
 * the pole route: D_B = 2 Re sum_j w_j (z_j S_W - H_W)^-1 [T,T] per window. */
#ifndef LRES_H
#define LRES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lres_plan lres_plan;

/* Precision of the pole solves. */
enum { LRES_FP64 = 0, LRES_MIXED = 1 };

/* Message of the last failed call on this thread ("" if none). */
const char* lres_last_error(void);

/* Number of HIP devices visible (0 when no GPU); never fails. */
int lres_device_count(void);

/* Window and ownership tables of a geometry (lrft/lcao3d.py:132-161).
 *   cells (ncell x 3), atom_pos (natom x 3), atom_of_fn (nb), box (3), grid shape (3).
 * Atoms go to their nearest cell centre (minimum image, lowest index on ties);
 * grid point (i,j,k) at (i,j,k)*dx goes to the centre with the smallest squared
 * minimum-image distance, lowest index on ties; W(B) lists, ascending, the
 * functions whose cell lies within R_w of B.  Outputs: h_Woff (ncell+1), and
 * with h_Wflat given, h_Wflat, h_Ooff (ncell+1), h_Oflat (ng); h_Wflat NULL
 * (a first call) fills h_Woff only, so that the caller can size h_Wflat. */
int lres_build_windows(const double* h_cells, int64_t ncell, const double* h_atom_pos, int64_t natom,
                       const int64_t* h_atom_of_fn, int64_t nb, const double* h_box, const int64_t* h_shape,
                       double dx, double R_w, int64_t* h_Wflat, int64_t* h_Woff, int64_t* h_Oflat,
                       int64_t* h_Ooff);

/* The same tables with the ng x ncell Voronoi distances on the GPU (the bulk of
 * lcao3d.py:141-151's cost: 2.6e8 distances at 1024 electrons); bitwise the
 * same output as lres_build_windows (identical operation order, no
 * contraction).  Needs a HIP device; at most 4096 cells. */
int lres_build_windows_device(const double* h_cells, int64_t ncell, const double* h_atom_pos, int64_t natom,
                              const int64_t* h_atom_of_fn, int64_t nb, const double* h_box,
                              const int64_t* h_shape, double dx, double R_w, int64_t* h_Wflat, int64_t* h_Woff,
                              int64_t* h_Oflat, int64_t* h_Ooff);

/* Per-geometry plan (replaces LRFTLCAOCpp.__init__'s flattening, lcao3d.py:343-357,
 * plus everything lrft_windows recomputes per call that only depends on the
 * geometry: T_B, Q_B, the owned rows of X, the dense S_W).  CSR arrays are
 * int64 indptr/indices + float64 values as produced by _csr64 (lcao3d.py:297-300).
 * M: number of circle poles (even).  precision: LRES_FP64 or LRES_MIXED.
 * stream: hipStream_t (NULL = null stream). */
int lres_plan_create(lres_plan** out, int64_t nb, int64_t ng, double dV, int M, int precision,
                     const int64_t* h_Wflat, const int64_t* h_Woff, int64_t nwin,
                     const int64_t* h_Oflat, const int64_t* h_Ooff,
                     const int64_t* h_Sp, const int64_t* h_Si, const double* h_Sv,
                     const int64_t* h_Xp, const int64_t* h_Xi, const double* h_Xv, void* stream);

int lres_plan_destroy(lres_plan* plan);

/* Window sizes n_W and |T_B| (functions with support on own(B)); h_n, h_k of length nwin. */
int lres_plan_sizes(const lres_plan* plan, int32_t* h_n, int32_t* h_k);

/* Window buckets of the plan (windows grouped by size, one pole-kernel launch
 * each): for bucket i < min(cap, *h_nbucket), h_out[8i..8i+7] = {largest n_W,
 * windows, fp64 T, fp64 G, fp64 kind, fp32 T, fp32 G, fp32 kind} (zeros until
 * the first pass in that precision has run the autotuner).  kind 0 / 1: vector
 * ALU Gauss-Jordan with a T x T tile per thread on a G x G grid and a register
 * budget of 128 / 512; kind 2: matrix cores (MFMA), T = 16 x 16 tiles per
 * dimension, G = waves per (window, pole group); G = 0 (complex64, chosen only
 * with LRES_POLE_FUSED=1 or =2) is the window-fused form: four waves per window, one
 * pole each, the pole sums and the reduction in the kernel, the same bits.  By default
 * buckets of up to 128 functions use kind 2 in both precisions and larger ones the
 * vector-ALU kernels; the environment variable LRES_POLE_KERNEL=valu, read at plan
 * creation and tuning, selects the vector-ALU family for both. */
int lres_plan_buckets(const lres_plan* plan, int32_t* h_out, int cap, int* h_nbucket);

/* Set precision (LRES_FP64 / LRES_MIXED) for later passes. */
int lres_plan_set_precision(lres_plan* plan, int precision);

/* Load H(v) as a host CSR (the Hc argument triple of lrft_windows,
 * src/lrft_cpp.cpp:129) and gather the dense window blocks H_W on the device. */
int lres_set_hamiltonian_csr(lres_plan* plan, const int64_t* h_Hp, const int64_t* h_Hi, const double* h_Hv,
                             int64_t nnz);

/* Same, with the CSR already in device memory (d_Hp has nb+1 entries). */
int lres_set_hamiltonian_csr_device(lres_plan* plan, const int64_t* d_Hp, const int64_t* d_Hi,
                                    const double* d_Hv);

/* Basis data for assembling H(v) on the device (replaces LCAOBasis.H,
 * lrft/lcao3d.py:98-100, and the host CSR round trip of lcao3d.py:411):
 * atom_of_fn (nb), X (grid x nb CSR, as at plan creation), the fixed kinetic
 * matrix T (nb x nb CSR, LCAOBasis.T, lcao3d.py:92-96) and the window tables of
 * plan creation.  At most 8 functions per atom. */
int lres_plan_set_basis(lres_plan* plan, const int64_t* h_atom_of_fn, int64_t natom,
                        const int64_t* h_Xp, const int64_t* h_Xi, const double* h_Xv,
                        const int64_t* h_Tp, const int64_t* h_Ti, const double* h_Tv,
                        const int64_t* h_Wflat, const int64_t* h_Woff);

/* H_W = T_W + [X^T diag(v) X dV]_W for every window from a device potential
 * d_v (ng doubles, C order of the grid); afterwards passes use this H.  Only
 * kernel launches on the plan's stream, so it can be captured into a graph. */
int lres_set_potential(lres_plan* plan, const double* d_v);

/* Make `stream` (hipStream_t, NULL = the null stream) the plan's stream for the
 * calls that follow, e.g. a stream being captured into a graph.  Ordering work
 * already queued on the previous stream is the caller's job.  No counterpart in
 * the reference (its pass is a blocking host call); for callers that capture
 * the SCF iteration. */
int lres_plan_set_stream(lres_plan* plan, void* stream);

/* The SCF driver's density mixing (lcao3d.py:268-281 in DIIS-coefficient form):
 * d_out = max(sum_h c_h (N_h + mix R_h), lo), then d_out *= ntot / (sum d_out * dV),
 * with c = d_coef (nhist device doubles), N_h / R_h rows h of d_Nh / d_Rh
 * (nhist x ng, the stored input densities and residuals).  The sum has a fixed
 * order; d_part is a device workspace of LRES_MIX_PARTS doubles, fully written,
 * so nothing needs clearing and the launches can be captured into a graph.
 * Kernel launches only, on `stream`. */
#define LRES_MIX_PARTS 1024
int lres_diis_mix(const double* d_coef, int nhist, const double* d_Nh, const double* d_Rh, int64_t ng, double mix,
                  double lo, double ntot, double dV, double* d_out, double* d_part, void* stream);

/* The SCF driver's potential v = v_ion + v_H + v_xc(n) (lcao3d.py:244-245) with
 * v_xc the Slater + Perdew-Zunger 1981 LDA potential of lda_xc (three_d.py:132-148,
 * density clamped at 1e-14); d_vh is the Hartree potential of d_n.  All device
 * arrays of ng doubles.  One kernel launch on `stream` (capturable). */
int lres_lda_potential(const double* d_n, const double* d_vion, const double* d_vh, int64_t ng, double* d_v,
                       void* stream);

/* The SCF driver's residual step (lcao3d.py:250-251, 258-265): r = scale * d_nout - d_nin,
 * stored as history row `slot` (d_Rh[slot] = r, d_Nh[slot] = d_nin; d_Nh / d_Rh are
 * nhist x ng), and d_stats[0] = sum |r|, d_stats[1 + q] = R_q . r for q < nhist (the
 * new row of DIIS's Gram matrix; rows not in use hold zeros).  nhist <= LRES_MAX_HIST.
 * Fixed-order sums; d_part is a device workspace of (LRES_MAX_HIST + 1) x LRES_MIX_PARTS
 * doubles.  d_stats may be pinned host memory (device-accessible): the totals are then
 * readable on the host once the stream has been waited for.  Kernel launches only, on
 * `stream`. */
#define LRES_MAX_HIST 8
int lres_scf_residual(const double* d_nout, double scale, const double* d_nin, int64_t ng, int nhist, int slot,
                      double* d_Nh, double* d_Rh, double* d_part, double* d_stats, void* stream);

/* lres_scf_residual without the d_part workspace argument: the same kernels, same
 * results bit for bit, with the partial sums in a static device array of
 * (LRES_MAX_HIST + 1) x LRES_MIX_PARTS doubles (72 KiB per GPU, held until the
 * process ends; its device address is looked up on the first call on each
 * device).  Allocates nothing (no hipMalloc / hipFree, no allocating library
 * call), for callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid it.  nhist above LRES_MAX_HIST is
 * rejected (the workspace's row limit).  Calls that overlap on one GPU (on
 * different streams, or from several threads) share the workspace and race;
 * order them on one stream.  No counterpart in the reference. */
int lres_scf_residual_noalloc(const double* d_nout, double scale, const double* d_nin, int64_t ng, int nhist,
                              int slot, double* d_Nh, double* d_Rh, double* d_stats, void* stream);

/* The integrals of the Harris-Foulkes energy at density d_n with Hartree potential
 * d_vh (lcao3d.py:276-279): d_out[0] = sum n v_H, d_out[1] = sum n e_xc(n),
 * d_out[2] = sum n (v_H + v_xc(n)) (grid sums, no dV).  d_part: 3 x LRES_MIX_PARTS
 * doubles of workspace.  Kernel launches only, on `stream`. */
int lres_harris_terms(const double* d_n, const double* d_vh, int64_t ng, double* d_part, double* d_out,
                      void* stream);

/* One window pass at (mu, e_lo): the pole form of lrft_windows
 * (src/lrft_cpp.cpp:129-233).  Per window B it forms
 *   D_B = 2 Re sum_{j<M/2} w_j G_j[T,T],  F_B = 2 Re sum_j w_j z_j G_j[T,T],
 *   G_j = (z_j S_W - H_W)^-1, z_j = c + R e^{i pi (2j+1)/M}, w_j = R e^{i pi (2j+1)/M} / M,
 * and writes N_B = Tr(D_B Q_B) (electrons of own(B), no spin factor),
 * E_B = Tr(F_B Q_B) and, if d_dens != NULL, d_dens[g] = x_g^T D_B x_g for every
 * owned grid point g (no spin factor, as lrft_windows' dens).
 * d_NB, d_EB: device arrays of nwin doubles (may be NULL).  h_sums (host, may
 * be NULL) receives {sum_B N_B, sum_B E_B} in a fixed order (bitwise reproducible). */
int lres_window_pass(lres_plan* plan, double mu, double e_lo, double* d_dens, double* d_NB, double* d_EB,
                     double* h_sums);

/* Owned-point density of the last pass (its D_B blocks are kept on the
 * device): the density output of lres_window_pass without repeating the solves. */
int lres_density_last(lres_plan* plan, double* d_dens);

/* Inertia counts: for each shift sigma_s (s < nshift, host array), the number of
 * window eigenvalues (H_W, S_W) below sigma_s summed over windows (h_total[s])
 * and per window (d_counts, nwin x nshift int32, may be NULL).  Sylvester's law
 * on an unpivoted LDL^T of sigma S_W - H_W in FP64. */
int lres_inertia(lres_plan* plan, const double* h_sigma, int nshift, int64_t* h_total, int32_t* d_counts);

/* Lowest window eigenvalue over all windows (min_B lambda_min(B), what the cold
 * path of lcao3d.py:450-451 takes from full spectra) by inertia bisection to
 * absolute tolerance tol. */
int lres_lambda_min(lres_plan* plan, double tol, double* h_lambda_min);

/* Chemical potential: root of 2 sum_B N_B(mu) - N by Brent's method on device
 * passes, bracket [lo, hi] with count(lo) <= 0 <= count(hi); f_lo / f_hi are the
 * counts already known at lo / hi (NaN: evaluated here).  Stops when the
 * bracket is below xtol (brentq's criterion) or |count| <= ftol.  Replaces
 * _solve_mu (lcao3d.py:387-398), which the eigen route evaluates in closed
 * form.  Each trial pass also writes its density to d_dens (may be NULL).
 * h_info (4 doubles, may be NULL) receives {passes, count at the returned mu,
 * 1 if the returned mu was the last pass (d_dens and sum E_B then belong to
 * it), sum_B E_B of that pass}. */
int lres_mu_solve(lres_plan* plan, double N, double e_lo, double lo, double hi, double f_lo, double f_hi,
                  double xtol, double ftol, double* d_dens, double* h_mu, double* h_info);

/* The warm step's mu search on the device side: from the previous mu0 (count f0
 * there, f = 2 sum N_B - N), secant steps with the last known slope (slope = 0:
 * steps of 1e-3 towards the root), each at most 0.25 and with 20% overshoot once
 * two points are on the same side, until |f| <= ftol or the root is bracketed,
 * then Brent (xtol, ftol) on f, or, when f(hi) > 10 |f(lo)|, on log(1 + f / A),
 * A = 1.25 |f(lo)| (the count grows exponentially towards the conduction band; the
 * transform keeps the root and straightens such a bracket); at most 60 secant steps.  No
 * counterpart in the reference (its mu search runs on stored eigenvalues); the same
 * steps as the host-side search it replaces, one call instead of one per pass.
 * h_info[8]: passes run, f at the returned mu, 1 if the last pass was at the
 * returned mu (its D_B blocks stay for lres_density_last), sum E_B there, the slope
 * after the search, 1 if no bracket was found (the returned mu is then NaN),
 * sum N_B there, 1 if the root came from Brent. */
int lres_mu_solve_warm(lres_plan* plan, double N, double e_lo, double mu0, double f0, double slope, double ftol,
                       double xtol, double* h_mu, double* h_info);

/* The warm density step of LRFTLCAOCpp._density (lcao3d.py:412-448: keep the filter
 * edge, keep mu, or re-solve mu, then the density) queued on the plan stream with the
 * decisions taken on the device, so that the host waits once per step instead of once
 * per trial pass:
 *   - the keep-edge test: inertia counts at e_lo + margin / 3 and e_lo + 2 margin
 *     (lcao3d.py:420; the edge is kept when the first is 0 and the second positive);
 *   - a pass at mu0, accepted when |count| <= keep_tol (or lag != 0 and |count| < 0.25,
 *     lcao3d.py:423-427);
 *   - else the search of lres_mu_solve_warm (same code, run by one device thread
 *     between passes) with slope, ftol, xtol;
 *   - the density of the accepted mu (no spin factor) into d_dens, from the last
 *     trial's D_B or from a pass of its own.
 * `depth` (1..8) pass slots are queued, the keep pass first; those the search does
 * not need leave at once, and lres_step_finish queues further batches (of at least 4)
 * when it needs more.  Work queued after this call on the same stream reads d_dens as
 * formed by the first batch (it holds the step's density unless a further batch ran,
 * which h_out reports); until lres_step_finish returns, no other call may use the plan. */
int lres_step_enqueue(lres_plan* plan, double N, double mu0, double e_lo, double margin, double keep_tol,
                      double ftol, double xtol, double slope, int lag, int depth, double* d_dens);

/* lres_step_enqueue in two halves, so that the step can be captured into a graph
 * (no counterpart in the reference: its warm step runs on the host, lcao3d.py:412-448):
 * lres_step_params writes the step's inputs (those of lres_step_enqueue but depth and
 * d_dens) to the plan's parameter block in pinned host memory, which the step's
 * kernels read when they run; lres_step_launch queues the first batch from it on the
 * plan's stream.  Launched while that stream is captured, the batch records nothing
 * host-dependent (no allocation, wait, memset or timing event; the kernels must have
 * been tuned by one launch outside a capture) and a replay of the graph runs the step
 * with the parameters written before it; lres_step_arm(depth) then tells the plan that
 * a replay queued a step of `depth` slots, and lres_step_finish waits for it as for an
 * enqueued one.  lres_step_enqueue = lres_step_params + lres_step_launch. */
int lres_step_params(lres_plan* plan, double N, double mu0, double e_lo, double margin, double keep_tol,
                     double ftol, double xtol, double slope, int lag);
int lres_step_launch(lres_plan* plan, int depth, double* d_dens);
int lres_step_arm(lres_plan* plan, int depth);

/* Waits for the queued step (queuing more trial batches when the search needs them)
 * and returns h_out[10] = {status (0: done, d_dens holds the density; 1: the keep-edge
 * test failed and nothing was computed; 2: no mu bracket within 60 secant steps),
 * outcome (0 kept mu, 1 lagged mu, 2 re-solved), passes run, mu, count at mu,
 * sum N_B and sum E_B at mu, slope after the search, 1 if the root came from Brent,
 * batches}.  A search that does not converge in 200 Brent steps is an error. */
int lres_step_finish(lres_plan* plan, double* d_dens, double* h_out);

/* One-shot form with lrft_windows' argument list (src/lrft_cpp.cpp:129-132):
 * host CSR H, S, X, window tables, nb, dV, mu, e_lo, M; outputs h_dens (ng, may be
 * NULL), h_NB, h_EB (nwin).  Builds and destroys a plan on the null stream. */
int lres_windows(const int64_t* h_Hp, const int64_t* h_Hi, const double* h_Hv,
                 const int64_t* h_Sp, const int64_t* h_Si, const double* h_Sv,
                 const int64_t* h_Xp, const int64_t* h_Xi, const double* h_Xv,
                 const int64_t* h_Wflat, const int64_t* h_Woff, int64_t nwin,
                 const int64_t* h_Oflat, const int64_t* h_Ooff, int64_t nb, int64_t ng, double dV,
                 double mu, double e_lo, int M, double* h_dens, double* h_NB, double* h_EB);

/* Cumulative pole-kernel statistics of passes that returned host sums:
 * h_out4 = {fp64 passes, fp64 kernel ms, fp32 passes, fp32 kernel ms}, measured
 * with HIP events around the pole-kernel launches on the plan's stream.
 * reset != 0 zeroes the counters after reading. */
int lres_plan_stats(lres_plan* plan, double* h_out4, int reset);

/* Keep-edge bound statistics (no counterpart in the reference: a diagnostic of how the
 * keep-edge test of lcao3d.py:420 was decided): h_out2 = {warm steps whose test was
 * checked against the bound, of them decided by the bound without inertia counts}.
 * The bound: with S = X^T X dV and H(v) = T + X^T diag(v) X dV, every window eigenvalue
 * moves by at most max |v - v_ref| from its value at the potential v_ref of the last
 * lres_lambda_min, so the test's outcome is proved when that bracket, widened by the
 * potential change, lies inside (e_lo + margin / 3, e_lo + 2 margin).  Cumulative over
 * the plan's life; LRES_EDGE_BOUND=0 in the environment switches the bound off. */
int lres_plan_edge_stats(lres_plan* plan, double* h_out2);

/* Timing of the last lres_window_pass, in milliseconds, measured with HIP events
 * on the plan's stream: {gather (set_hamiltonian), pole solves, reduction}. */
int lres_last_timing(const lres_plan* plan, double* h_ms3);

/* Layout of the plan's per-window blocks (H_W, S_W and the adjoint's
 * gradients): window w occupies doubles [h_mat_off[w], h_mat_off[w + 1]) as an
 * n_W x n_W row-major matrix in "T-last" order; h_fn (optional, length
 * sum n_W, window w from Woff[w] - Woff[0]) gives the global basis function at
 * each position.  Replaces the window lists of LRFTLCAO.__init__
 * (lrft/lcao3d.py:155) for callers that map block gradients back to H and S. */
int lres_plan_layout(const lres_plan* plan, int64_t* h_mat_off, int64_t* h_fn);

/* Replace the plan's window blocks H_W and/or S_W (device, the lres_plan_layout
 * layout; NULL keeps a block), or copy them out.  For differentiating through
 * the window step; no counterpart in the reference. */
int lres_set_window_blocks(lres_plan* plan, const double* d_HW, const double* d_SW);
int lres_get_window_blocks(const lres_plan* plan, double* d_HW, double* d_SW);

/* D_B and F_B of the last pass (device, sum_B k_B^2 doubles: window after
 * window, k_B x k_B row-major over T_B in ascending function order, which is
 * positions n_W - k_B .. n_W - 1 of the lres_plan_layout order):
 * D_B = 2 Re sum_j w_j G_j[T,T], F_B = 2 Re sum_j w_j z_j G_j[T,T].  The
 * density-matrix blocks behind lrft/lcao3d.py:126-129's counts; with the
 * adjoint they give the gradient with respect to Q_B (d E_B / d Q_B = F_B,
 * d N_B / d Q_B = D_B). */
int lres_pass_blocks(const lres_plan* plan, double* d_D, double* d_F);

/* Adjoint of the window step at the fixed circle (mu, e_lo), fp64: the
 * gradients of sum_B (alpha E_B + beta N_B) with respect to every window's H_W
 * and S_W, written to d_gH and d_gS (device, lres_plan_layout layout, no spin
 * factor):
 *   dH_W:  2 Re sum_j w_j p(z_j)     G_j[:,T] Q_B G_j[T,:]
 *   dS_W: -2 Re sum_j w_j z_j p(z_j) G_j[:,T] Q_B G_j[T,:],   p(z) = alpha z + beta,
 * G_j = (z_j S_W - H_W)^-1 (lrft/core.py:130-139 poles; N_B, E_B as
 * lres_window_pass).  alpha = 1, beta = 0 gives the band-energy gradient;
 * alpha = 0, beta = 1 the count's.  The fixed-N derivative of the band energy
 * (mu moving with the perturbation, e_lo fixed) is dE - (dE/dmu)/(dN/dmu) dN,
 * i.e. alpha = 1, beta = -(dE/dmu)/(dN/dmu).  No counterpart in the reference
 * (its 3D path has no forces, README.md:35).
 * Allocates workspaces of up to 2 x 4 GiB (windows are processed in chunks). */
int lres_adjoint(lres_plan* plan, double mu, double e_lo, double alpha, double beta, double* d_gH, double* d_gS);

/* Per-geometry set-up of the device force contraction: the basis functions of every
 * atom (the Gaussians s_exps[0..ns) then x, y, z per p exponent, times
 * (1 - (r/rcut)^2)^2, normalised by the grid quadrature: LCAOBasis,
 * lrft/lcao3d.py:49-100) and their derivatives with respect to the atom's position,
 * tabulated on a dense box of (2 ceil(rcut/dx) + 1)^3 grid points around round(R/dx)
 * (h_pos natom x 3, h_box the cell lengths, h_shape the grid); the atom pairs whose
 * boxes meet, with their entries in the plan's window blocks; the owning window of
 * every grid point.  Functions must be atom-major with ns + 3 np (<= 8) per atom.
 * Device memory: 8 (5 nf) (2 ceil(rcut/dx) + 1)^3 bytes per atom.  No counterpart in
 * the reference (its 3D path has no forces, README.md:35). */
int lres_force_setup(lres_plan* plan, const double* h_pos, int64_t natom, const int64_t* h_atom_of_fn, int64_t nb,
                     double rcut, const double* h_exps, int ns, int np, const double* h_box, const int64_t* h_shape,
                     double dx);

/* Band-energy forces F_A = -dE/dR_A (d_forces, natom x 3 device doubles) from the
 * window step's gradients: d_gH, d_gS the lres_adjoint output (layout of
 * lres_plan_layout) for alpha = 1 and the given beta, d_D, d_F the lres_pass_blocks
 * output of the pass at the same (mu, e_lo), d_v the grid potential of H (ng device
 * doubles).  With the spin factor 2, G^H = 2 sum_W d_gH, G^S = 2 sum_W d_gS and
 * G^Q_B = 2 (F_B + beta D_B):
 *   dE/dR_A = 2 dV sum_{c on A} sum_g dx_c/dR_A [ sum_b G^H_cb ((T + v) x_b)
 *                                                 + G^S_cb x_b + sum_u G^Q_{B(g),cu} x_u ](g)
 * (H = X^T (T + diag v) X dV, S = X^T X dV, Q_B = sum_{g in own B} x_g x_g^T dV, T the
 * 7-point kinetic operator of three_d.py:151-164).  Fixed-order sums: bitwise
 * reproducible.  Kernel launches only, on the plan's stream. */
int lres_band_forces(lres_plan* plan, const double* d_gH, const double* d_gS, const double* d_D, const double* d_F,
                     double beta, const double* d_v, double* d_forces);

#ifdef __cplusplus
}
#endif

#endif /* LRES_H */
