// Internal declarations shared by the kernel and host translation units of liblres_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lres_musearch.h"

namespace lres {

// One window of a plan (all offsets in elements).
struct WinDesc {
    int64_t mat_off;   // n x n dense blocks in SW / HW
    int64_t kk_off;    // k x k blocks in Q / D / F (times the pole-group count in Dpart / Fpart)
    int64_t own_off;   // first owned point in the owned-row CSR
    int32_t n, k, m, nown;
};

// Outputs of a pole kernel that also does the reduction of its windows (the complex64
// matrix-core kernel with all poles of a window in one workgroup): D_B, F_B (upper
// triangles, as k_reduce writes them), Q_B read, N_B and E_B written.
struct WinOut {
    double* D;
    double* F;
    const double* Q;
    double* NB;
    double* EB;
};

}  // namespace lres

hipError_t launch_gather_csr(const lres::WinDesc* wins, int nwin, const int64_t* Wflat, const int64_t* Woff,
                             const int32_t* pos, const int64_t* Hp, const int64_t* Hi, const double* Hv,
                             double* out, hipStream_t st);
// The pole kernels read the circle of their pass from device memory: circ = {c, R,
// skip, skip_density}; skip != 0 makes every workgroup leave at once (a trial pass the
// device-side mu search did not need), skip_density != 0 does the same for the density
// kernel.  k_set_circle writes it for a pass decided on the host.
hipError_t launch_set_circle(double* circ, double c, double R, double skip, double skip_density, hipStream_t st);
// The device side of the queued warm density step (lres_step_enqueue): k_mu_begin
// starts the search of lres_musearch.h (keep test at mu0, after the keep-edge test on
// the inertia totals itot if given); k_sum_step sums a slot's N_B and E_B into out and
// runs one controller stage (StepStage) on them; kStepPublish copies the state to `pin`
// (mapped host memory); k_mu_step runs a stage without sums (kStepResume).
// The inputs of a device step, in this order, in a parameter block (doubles) the host
// writes before launching the step or replaying a graph that holds it.
enum StepPrm {
    kStepPrmN = 0, kStepPrmMu0, kStepPrmElo, kStepPrmMargin, kStepPrmKeepTol, kStepPrmFtol, kStepPrmXtol,
    kStepPrmSlope, kStepPrmLag, kStepPrmCount
};
hipError_t launch_mu_begin(lres::StepState* st, const unsigned long long* itot, const double* prm, double* circ,
                           hipStream_t stream);
hipError_t launch_mu_step(lres::StepState* st, int stage, double* circ, void* pin, hipStream_t stream);
// Keep-edge bound (lres_kernels.hip): edge = {lo, hi, delta, skipped, checked}.
// launch_edge_delta copies v to vcur and sets delta = max |v - vref| (kEdgeBlocks
// partial maxima in part); launch_edge_ref starts a reference bracket [lo, hi] (delta
// 0); the keep-edge test of a device step is decided from it by launch_step_edge.
constexpr int kEdgeBlocks = 256;
// keep-edge bound state: [0, 1] the reference bracket of the smallest window eigenvalue,
// [2] the largest rise and [5] the largest fall of the potential since, [3] tests the
// bound decided, [4] tests checked
constexpr int kEdgeState = 6;
// lambda_min on the device (lres_kernels.hip): the search state, read by the count
// launches of a round through &ncand ({ncand, K, done}); kLminProbe bracket probes below
// the Rayleigh bound, about kLminWaves (window, shift) factorisations per round (one
// wave each: about two per SIMD, one generation of waves, so a round costs about one
// factorisation's latency), at least kLminMinK and at most kLminMaxK shifts; the host
// search uses the same rule.
struct LminState {
    double lo, hi, tol;
    double cell;                // the dyadic cell [cell, cell + width) holding lambda_min (lmin_snap)
    int ncand, K, done, fail;   // fail: 1 no count at the Rayleigh bound, 2 no probe counts zero
};
// The result does not depend on the path of the search: the final bracket [lo, hi)
// (hi - lo <= tol) is replaced by the cell of the dyadic grid of the given width (the
// smallest power of two >= tol) that holds lambda_min, which needs the count at the
// grid point g inside the bracket when there is one.  lmin_grid: g, or NAN when the
// bracket lies in one cell (*cell is then set); exact arithmetic (powers of two).
__host__ __device__ inline double lmin_grid(double lo, double hi, double width, double* cell) {
    const double g = ceil(lo / width) * width;
    if (lo == g) {
        *cell = g;
        return NAN;
    }
    if (hi <= g) {
        *cell = g - width;
        return NAN;
    }
    return g;
}
constexpr int kLminProbe = 2, kLminWaves = 2048, kLminMinK = 3, kLminMaxK = 255;
hipError_t launch_lmin_probe(const double* rq, int nwin, LminState* s, double tol, double* sig,
                             unsigned long long* itot, hipStream_t st);
hipError_t launch_lmin_bracket(LminState* s, double* sig, unsigned long long* itot, const int32_t* ids, int nwin,
                               int32_t* cand, hipStream_t st);
hipError_t launch_lmin_round(LminState* s, double* sig, unsigned long long* itot, const int32_t* counts, int kcap,
                             int32_t* cand, int32_t* tmp, hipStream_t st);
// the snap of a finished search: launch_lmin_snap sets sig[0] = g and *skip = 0 when the
// count at g is needed (else *skip = 1 and the cell is set), launch_lmin_snap_end takes
// the cell from that count (itot[0])
hipError_t launch_lmin_snap(LminState* s, double width, double* sig, unsigned long long* itot, int* skip,
                            hipStream_t st);
hipError_t launch_lmin_snap_end(LminState* s, double width, const double* sig, const unsigned long long* itot,
                                const int* skip, hipStream_t st);
hipError_t launch_edge_delta(const double* v, double* vcur, const double* vref, int64_t ng, double* part,
                             double* edge, hipStream_t st);
hipError_t launch_edge_ref(double* edge, double lo, double hi, hipStream_t st);
// the keep-edge test's set-up for a device step (shifts, zeroed totals, the bound's
// decision when edge is given), from the parameter block prm
hipError_t launch_step_edge(const double* prm, double* sig, unsigned long long* itot, double* edge, double safety,
                            int* skip, hipStream_t st);
hipError_t launch_sum_step(const double* a, const double* b, int n, double* out, lres::StepState* st, int stage,
                           double* circ, void* pin, hipStream_t stream);
hipError_t launch_pole_gj(int T, int G, int wide, int fp32, const lres::WinDesc* wins, const int32_t* ids, int nids,
                          const double* SW, const double* HW, int M, int ngroup, const double* circ,
                          double* Dpart, double* Fpart, hipStream_t st);
// D, F, N_B, E_B from the pole kernels' partials: group sums (Dpart, Fpart) or,
// for windows with by_pole[w] == 1, nslot per-pole float pairs in `slots`; windows with
// by_pole[w] == 2 were reduced by the window-fused pole kernel and are left alone
hipError_t launch_reduce(const lres::WinDesc* wins, int nwin, int ngroup, const double* Dpart, const double* Fpart,
                         const float* slots, int nslot, const int8_t* by_pole, const double* Q, double* D, double* F,
                         double* NB, double* EB, const double* circ, hipStream_t st);
// lower triangles of D_B (and F_B if F != NULL) from the upper ones the reduction writes
hipError_t launch_mirror(const lres::WinDesc* wins, int nwin, double* D, double* F, hipStream_t st);
// The owned rows of X restricted to T_B (CSR over owned points, T-local int32 columns)
// and, when every row is whole 6-function atom blocks, the same rows in 64-point
// chunks, block-interleaved (chk_off null: no chunked form; the row-per-wave kernel).
struct DensRows {
    const int64_t* own_g;
    const int64_t* xptr;
    const int32_t* xidx;
    const double* xval;
    const int64_t* chk_off;
    const int64_t* chk_bs;
    const int32_t* chk_idx;
    const double* chk_val;
};
// skip (may be null): the density kernel leaves at once when *skip != 0
hipError_t launch_density(const lres::WinDesc* wins, int nwin, int kmax, const double* D, DensRows rows, double* dens,
                          const double* skip, hipStream_t st);
hipError_t launch_sum2(const double* a, const double* b, int n, double* out, const double* skip, hipStream_t st);
hipError_t launch_voronoi_owner(const double* cells, int ncell, const double* box, const int64_t* shape, double dx,
                                int32_t* owner, hipStream_t st);
// SCF density mixing: grid and block counts of its fixed-order sum (kMixBlocks partials)
constexpr int kMixBlocks = 1024, kMixThreads = 256;
hipError_t launch_diis_mix(const double* coef, int nh, const double* Nh, const double* Rh, int64_t ng, double mix,
                           double lo, double ntot, double dV, double* out, double* part, hipStream_t st);
// SCF grid passes: LDA potential, residual + DIIS dot products (history rows <= kMaxHist),
// Harris-Foulkes integrals; sums over kMixBlocks partials per quantity
constexpr int kMaxHist = 8;
hipError_t launch_lda_potential(const double* n, const double* vion, const double* vh, int64_t ng, double* v,
                                hipStream_t st);
hipError_t launch_scf_residual(const double* nout, double scale, const double* nin, int64_t ng, int nh, int slot,
                               double* Nh, double* Rh, double* part, double* stats, hipStream_t st);
hipError_t launch_harris(const double* n, const double* vh, int64_t ng, double* part, double* out, hipStream_t st);
hipError_t launch_diag_ratio(const lres::WinDesc* wins, int nwin, const double* SW, const double* HW, double* out,
                             hipStream_t st);
hipError_t launch_inertia(int T, int G, const lres::WinDesc* wins, const int32_t* ids, int nids, const double* SW,
                          const double* HW, const double* sigma, int nshift, int32_t* counts,
                          unsigned long long* total, const int* skip, hipStream_t st);
hipError_t launch_pair_v(int nf_max, int npair, const int32_t* pairA, const int32_t* pairB, const int64_t* pent_off,
                         const int32_t* ent_r, const int32_t* ent_a, const int32_t* ent_b, const int64_t* xat_off,
                         const int32_t* nf_of, const double* xat, const int64_t* vp_off, const double* v, double dV,
                         double* vpair, hipStream_t st);
// HW[q] = TW[q] + vpair[hmap[q]] (TW[q] where hmap[q] < 0) over total entries
hipError_t launch_window_hv(int64_t total, const int32_t* hmap, const double* vpair, const double* TW, double* HW,
                            hipStream_t st);
// Windows above the register-resident limit (blocked sweep in a global workspace).
size_t big_lds_bytes(int nmax, int b, int elem_bytes);
hipError_t launch_pole_big(int fp32, int bmax, int nmax, const lres::WinDesc* wins, const int32_t* ids, int nids,
                           const double* SW, const double* HW, int M, int ngroup, const double* circ, void* work,
                           const int64_t* ws_off, double* Dpart, double* Fpart, hipStream_t st);
// Inertia counts of windows above the register-resident limit: work holds work_len
// doubles, cut into slots of nmax^2; up to work_len / nmax^2 (window, shift) pairs are
// factored at once (at least nids slots, else hipErrorInvalidValue).
hipError_t launch_inertia_big(int bmax, int nmax, const lres::WinDesc* wins, const int32_t* ids, int nids,
                              const double* SW, const double* HW, const double* sigma, int nshift, double* work,
                              int64_t work_len, int32_t* counts, unsigned long long* total, const int* skip,
                              hipStream_t st);
// Pole solves on the matrix cores for windows of up to 128 functions: nt = 16 x 16
// tiles per dimension (2..8), nwc = waves per (window, pole group) splitting the
// tile columns (1..nt).  One object per (precision, nt),
// compiled in parallel.
// nwc > 0: waves per (window, pole group); nwc == 0 (fp32 only): the window-fused form,
// four waves per window each on its own pole, the poles' D and F summed in LDS in pole
// order and reduced in place into `out` (no partials, no k_reduce for those windows).
#define LRES_MFMA_PART_DECL(r, nt)                                                                              \
    hipError_t lres_mfma_part_##r##_##nt(int nwc, const lres::WinDesc* wins, const int32_t* ids, int nids,     \
                                         const double* SW, const double* HW, int M, int ngroup,               \
                                         const double* circ, double* Dpart, double* Fpart, int acc,           \
                                         lres::WinOut out, hipStream_t st);
#define LRES_MFMA_PARTS(X) X(float, 2) X(float, 3) X(float, 4) X(float, 5) X(float, 6) X(float, 7) X(float, 8) \
    X(float, 9) X(float, 10) X(float, 11) X(float, 12) X(float, 13) X(float, 14) X(float, 15) X(float, 16)     \
    X(double, 2) X(double, 3) X(double, 4) X(double, 5) X(double, 6) X(double, 7) X(double, 8)
LRES_MFMA_PARTS(LRES_MFMA_PART_DECL)
#define LRES_MFMA_INERTIA_DECL(nt)                                                                              \
    hipError_t lres_mfma_inertia_##nt(const lres::WinDesc* wins, const int32_t* ids, int nids, const double* SW,   \
                                      const double* HW, const double* sigma, int nshift, int32_t* counts,       \
                                      unsigned long long* total, const int* skip, const int* lin, hipStream_t st);
#define LRES_MFMA_INERTIA_PARTS(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8)
LRES_MFMA_INERTIA_PARTS(LRES_MFMA_INERTIA_DECL)
// inertia counts on the matrix cores for windows of 16 nt functions at most (nt in 2..8)
// lin (may be null): a linear launch of nids workgroups over lin = {ncand, K, done}:
// workgroup i counts window ids[i / K] at shift i % K (none once done or past ncand),
// per-window counts at counts[w * nshift + shift]
hipError_t launch_inertia_mfma(int nt, const lres::WinDesc* wins, const int32_t* ids, int nids, const double* SW,
                               const double* HW, const double* sigma, int nshift, int32_t* counts,
                               unsigned long long* total, const int* skip, hipStream_t st,
                               const int* lin = nullptr);
// acc > 0 (fp64): the group partials are accumulated over the group's poles in LDS, as
// packed upper triangles of acc entries (at least k(k+1)/2 of every window launched)
hipError_t launch_pole_mfma(int fp32, int nt, int nwc, const lres::WinDesc* wins, const int32_t* ids, int nids,
                            const double* SW, const double* HW, int M, int ngroup, const double* circ,
                            double* Dpart, double* Fpart, int acc, lres::WinOut out, hipStream_t st);
// accumulator registers per lane of one wave of the MFMA pole kernel
int mfma_acc_regs(int fp32, int nt, int nwc);
// scratch bytes per lane of an MFMA variant (register spills; -1: no such variant)
int mfma_scratch_bytes(int fp32, int nt, int nwc);
// Adjoint at a fixed circle: d/dH_W and d/dS_W of sum_B (alpha E_B + beta N_B) for
// the windows ids[0..nids) (n <= nmax), fp64.  Workspaces Gw, Zw: woff[i] is the
// offset (complex entries) of window ids[i], (M/2) n^2 entries each.
hipError_t launch_adjoint(int bmax, int nmax, const lres::WinDesc* wins, const int32_t* ids, const int64_t* woff,
                          int nids, const double* SW, const double* HW, const double* Q, int M, double c, double Rr,
                          double alpha, double beta, void* Gw, void* Zw, double* gH, double* gS, hipStream_t st);
// Forces on the device (lres_forces.hip): the basis on per-atom boxes, the gradient
// blocks of the force pairs, the pair and Q parts and the per-atom totals.
hipError_t launch_atom_boxes(int natom, const double* pos, const int64_t* origin, int nbox, double dx,
                             const double* L, double rcut, const double* exps, int ns, int np, double dV, double* X,
                             double* DX, double* TX, hipStream_t st);
hipError_t launch_force_gather(int npair, int nf, const int64_t* ent_ptr, const int64_t* ent_off,
                               const int32_t* ent_n, const double* gH, const double* gS, double* GH, double* GS,
                               hipStream_t st);
hipError_t launch_force_pairs(int npair, int nf, const int32_t* pA, const int32_t* pB, const int64_t* origin, int nbox,
                              const int64_t* shape, const double* X, const double* DX, const double* TX,
                              const double* v, const double* GH, const double* GS, double* out, hipStream_t st);
hipError_t launch_force_q(int natom, int nf, const int64_t* origin, int nbox, const int64_t* shape, const double* X,
                          const double* DX, const int32_t* owner, const int32_t* row_of, const int32_t* tpos,
                          const lres::WinDesc* wins, const int64_t* xptr, const int32_t* xidx, const double* xval,
                          const double* D, const double* F, double beta, double* out, hipStream_t st);
hipError_t launch_force_total(int natom, const int64_t* aptr, const int64_t* aent, const double* fpair,
                              const double* fq, double dV, double* F, hipStream_t st);
