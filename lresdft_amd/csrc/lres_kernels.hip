// Device kernels of the LRes window step for CDNA4 (gfx950).
//
// Data layout in HBM (per plan):
//   SW, HW   dense window blocks, float64, row-major n x n, rows/columns in the
//            "T-last" order: the m = n - k window functions without support on
//            own(B) first, then the k functions of T_B (both ascending).
//   Q        per window k x k float64, Q_B = sum_{g in own(B)} x_g x_g^T dV.
//   Dpart/Fpart  per (window, pole group) k x k float64 partial pole sums, upper
//            triangle only (G[T,T] of a complex symmetric pencil is symmetric).
//   D/F      per window k x k float64: D_B = 2 Re sum_j w_j G_j[T,T], F_B likewise with w_j z_j.
//   owned rows of X restricted to T: CSR over owned points, int32 T-local columns.
//
// The pole solve of one (window, pole) is an in-place Gauss-Jordan sweep of
// A = z S_W - H_W held in registers: a 16 x 16 thread grid, thread (ty, tx)
// owning rows ty + 16 i and columns tx + 16 j (i, j < TS).  Pivots 0..m-1 only
// form the Schur complement on T (the rows of G outside T are never needed);
// pivots m..n-1 invert it, so the T x T block ends as (A^-1)[T,T].  No pivoting:
// Im(A) = Im(z) S_W is positive definite for Im z > 0, so every leading block
// is nonsingular.  Each step broadcasts the pivot row and column through LDS
// (double-buffered: one barrier per step) and applies a rank-1 update to the
// register tile.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "lres_device.h"
#include "lres_internal.h"

namespace lres {

// ---------------------------------------------------------------------------
// gather of dense H_W from a CSR H (src/lrft_cpp.cpp:117-127), one block per window
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gather_csr(const WinDesc* __restrict__ wins,
                                                    const int64_t* __restrict__ Wflat,
                                                    const int64_t* __restrict__ Woff,
                                                    const int32_t* __restrict__ pos_of_local,
                                                    const int64_t* __restrict__ Hp,
                                                    const int64_t* __restrict__ Hi,
                                                    const double* __restrict__ Hv, double* __restrict__ out) {
    const int w = blockIdx.x;
    const WinDesc d = wins[w];
    const int n = d.n;
    const int64_t* W = Wflat + Woff[w];
    const int32_t* pos = pos_of_local + Woff[w];
    double* A = out + d.mat_off;
    for (int64_t q = threadIdx.x; q < (int64_t)n * n; q += blockDim.x) A[q] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (int a = wave; a < n; a += nwave) {
        const int64_t g = W[a];
        const int pa = pos[a];
        for (int64_t p = Hp[g] + lane; p < Hp[g + 1]; p += 64) {
            const int64_t col = Hi[p];
            int lo = 0, hi = n - 1, hit = -1;
            while (lo <= hi) {   // W is ascending
                const int mid = (lo + hi) >> 1;
                const int64_t v = W[mid];
                if (v == col) { hit = mid; break; }
                if (v < col) lo = mid + 1; else hi = mid - 1;
            }
            if (hit >= 0) A[(int64_t)pa * n + pos[hit]] = Hv[p];
        }
    }
}

// ---------------------------------------------------------------------------
// pole solves: register-tiled Gauss-Jordan on a G x G thread grid, T x T tile
// per thread (cyclic: thread (ty, tx) owns rows ty + G i, columns tx + G j).
// G = blockDim.x / G is chosen per bucket so that G T >= n with little padding
// (n = 84: G = 21, T = 4, no padding, ~110 VGPRs -> 4 waves per SIMD).
// ---------------------------------------------------------------------------
constexpr int kMaxG = 32;

// Gauss-Jordan sweep without fix-up branches.  For pivot p with value a = A[p][p]
// take the pivot row with its pivot entry replaced by a + 1 and the pivot
// column with its pivot entry replaced by a - 1; then the plain rank-1 update
// A -= col (row / a) gives exactly the in-place Gauss-Jordan step: row p ->
// row p / a, column p -> -column p / a, A[p][p] -> 1 / a, all other entries
// A - col row / a.  (Proof: A[i][p] - A[i][p] (a+1)/a = -A[i][p]/a;
// A[p][j] - (a-1) A[p][j]/a = A[p][j]/a; a - (a-1)(a+1)/a = 1/a.)  The same
// update serves the Schur pivots 0..m-1: after all n pivots the T x T block
// holds (A^-1)[T, T] either way.
//
// A is complex symmetric, and after the pivots of a set D the in-place state is
// symmetric on D x D and on the rest U x U and antisymmetric between them
// (W[u][d] = -W[d][u]).  So only the entries on and above the diagonal are
// updated and only the pivot row is published: column p is read from it, with
// the sign flipped on the rows already pivoted.

// x - c r with the column factor passed pre-negated: nc = -c, ci = c.im
template <typename R>
__device__ __forceinline__ cx<R> cupd(cx<R> x, R ncr, R nci, R ci, cx<R> r) {
    R re = x.re, im = x.im;
    re = fma(ncr, r.re, re);
    re = fma(ci, r.im, re);
    im = fma(ncr, r.im, im);
    im = fma(nci, r.re, im);
    return {re, im};
}

// Publish the row of pivot q (tile slot PQ, compile-time after unrolling) from the
// owners' registers: the entries right of the diagonal (columns >= q) from thread
// row q % G, those left of it from thread column q % G as -A[j][q] (rows j < q are
// pivoted, q is not, so A[q][j] = -A[j][q]).  Raw values: the +1 / -1 shifts of
// the pivot entry are applied where the row is read.
template <int T, typename R>
__device__ __forceinline__ void publish_pivot(const cx<R> (&A)[T][T], int PQ, int q, bool own_row, bool own_col,
                                              int G, int tx, int ty, cx<R>* rowq) {
    if (own_row) {
#pragma unroll
        for (int j = 0; j < T; ++j) {
            const int col = tx + G * j;
            if (col >= q) rowq[col] = A[PQ][j];
        }
    }
    if (own_col) {
#pragma unroll
        for (int i = 0; i < T; ++i) {
            const int row = ty + G * i;
            if (row < q) rowq[row] = {-A[i][PQ].re, -A[i][PQ].im};
        }
    }
}

template <int T, int MAXT, typename R>
__global__ void __launch_bounds__(MAXT) k_pole_gj(const WinDesc* __restrict__ wins, const int32_t* __restrict__ ids,
                                                  const double* __restrict__ SW, const double* __restrict__ HW,
                                                  int M, int ngroup, int G, const double* __restrict__ circ,
                                                  double* __restrict__ Dpart, double* __restrict__ Fpart) {
    __shared__ cx<R> rowbuf[2][kMaxG * T];
    __shared__ cx<R> invbuf[2];
    if (circ[2] != 0.0) return;   // a trial pass the device-side mu search did not need
    const double c = circ[0], Rr = circ[1];
    const int w = ids[blockIdx.x];
    const int g = blockIdx.y;
    const WinDesc d = wins[w];
    const int n = d.n, m = d.m, k = d.k;
    const int ty = threadIdx.x / G, tx = threadIdx.x - ty * G;
    const double* S = SW + d.mat_off;
    const double* H = HW + d.mat_off;
    // this block's slot of the pole-group partial sums (each thread owns its tile entries)
    double* dp = Dpart + d.kk_off * ngroup + (int64_t)g * k * k;
    double* fp = Fpart + d.kk_off * ngroup + (int64_t)g * k * k;

    cx<R> A[T][T];
    bool first = true;
    for (int jp = g; jp < M / 2; jp += ngroup) {
        cx<double> z, wt, wz;
        pole(jp, M, c, Rr, z, wt, wz);
        // opaque per pole: keeps the compiler from hoisting T*T tile addresses out of the pole loop
        const double* Sp = S;
        const double* Hq = H;
        asm volatile("" : "+s"(Sp), "+s"(Hq));
        // A = z S - H (identity on the padding), one tile row of loads in flight at a time
#pragma unroll
        for (int i = 0; i < T; ++i) {
            const int r = ty + G * i;
#pragma unroll
            for (int j = 0; j < T; ++j) {
                if (j < i) continue;   // never used (below the diagonal)
                const int col = tx + G * j;
                const bool in = r < n && col < n;
                const double h = in ? Hq[(int64_t)r * n + col] : (r == col ? -1.0 : 0.0);
                const double sv = in ? Sp[(int64_t)r * n + col] : 0.0;
                A[i][j] = {(R)(z.re * sv - h), (R)(z.im * sv)};
            }
            asm volatile("" ::: "memory");
        }
        publish_pivot<T, R>(A, 0, 0, ty == 0, tx == 0, G, tx, ty, rowbuf[0]);
        if (ty == 0 && tx == 0) invbuf[0] = crecip(A[0][0]);
        // pivot p = G PI + pr: the tile slot PI is a compile-time constant in each copy
#pragma unroll
        for (int PI = 0; PI < T; ++PI) {
            const int pr_end = min(G, n - G * PI);
            for (int pr = 0; pr < pr_end; ++pr) {
                const int p = G * PI + pr, buf = p & 1;
                __syncthreads();
                const cx<R> inv = invbuf[buf];
                cx<R> rs[T];
                R ncr[T], nci[T], ci[T];
#pragma unroll
                for (int j = 0; j < T; ++j) {
                    const int col = tx + G * j;
                    cx<R> v = rowbuf[buf][col];
                    if (col == p) v.re += R(1);
                    rs[j] = cmul(v, inv);
                }
                // column p from the row: A[r][p] = -A[p][r] on the pivoted rows r < p
#pragma unroll
                for (int i = 0; i < T; ++i) {
                    const int r = ty + G * i;
                    cx<R> cv = rowbuf[buf][r];
                    if (r < p) cv = {-cv.re, -cv.im};
                    if (r == p) cv.re -= R(1);
                    ncr[i] = -cv.re;
                    nci[i] = -cv.im;
                    ci[i] = cv.im;
                }
                const int q = p + 1;
                const int qr = q - G * (PI + (pr + 1 == G ? 1 : 0));   // q % G
                const bool pub = q < n;
                const bool own_row = pub && ty == qr, own_col = pub && tx == qr;
                // the owner of the next pivot forms it and its reciprocal first (same
                // operations as in the bulk update below), off the other waves' path
                if (own_row && own_col) {
                    const int PQ = pr + 1 < G ? PI : PI + 1;
                    cx<R> a = {R(0), R(0)};
#pragma unroll
                    for (int s = 0; s < T; ++s)
                        if (s == PQ) a = cupd(A[s][s], ncr[s], nci[s], ci[s], rs[s]);
                    invbuf[buf ^ 1] = crecip(a);
                }
                // rank-1 update of the tile entries i <= j: every entry on or above the
                // diagonal that this thread owns is one of them (an entry with i > j has
                // row - col >= G (i - j) - (G - 1) > 0)
#pragma unroll
                for (int i = 0; i < T; ++i)
#pragma unroll
                    for (int j = 0; j < T; ++j)
                        if (j >= i) A[i][j] = cupd(A[i][j], ncr[i], nci[i], ci[i], rs[j]);
                if (pr + 1 < G)
                    publish_pivot<T, R>(A, PI, q, own_row, own_col, G, tx, ty, rowbuf[buf ^ 1]);
                else
                    publish_pivot<T, R>(A, PI + 1 < T ? PI + 1 : T - 1, q, own_row, own_col, G, tx, ty,
                                        rowbuf[buf ^ 1]);
            }
        }
        __syncthreads();   // the next pole's prologue rewrites buffer 0
        // accumulate Re(w G) and Re(w z G) on T x T into this block's partial slot;
        // pointers and sizes are made opaque per pole so that the per-entry
        // addresses are recomputed here instead of being held across the sweep
        double* dpp = dp;
        double* fpp = fp;
        int kk = k, mm = m;
        asm volatile("" : "+s"(dpp), "+s"(fpp), "+s"(kk), "+s"(mm));
#pragma unroll
        for (int i = 0; i < T; ++i) {
            const int r = ty + G * i - mm;
#pragma unroll
            for (int j = 0; j < T; ++j) {
                if (j < i) continue;   // below the diagonal
                const int col = tx + G * j - mm;
                if (r >= 0 && r <= col && col < kk) {
                    const double gr = (double)A[i][j].re, gi = (double)A[i][j].im;
                    const double dv = wt.re * gr - wt.im * gi, fv = wz.re * gr - wz.im * gi;
                    const int64_t q = (int64_t)r * kk + col;
                    if (first) {
                        dpp[q] = dv;
                        fpp[q] = fv;
                    } else {
                        dpp[q] += dv;
                        fpp[q] += fv;
                    }
                }
            }
        }
        first = false;
    }
}

// ---------------------------------------------------------------------------
// reduction over pole groups, N_B / E_B traces and the owned-point density
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double* red) {
    // fixed-shape tree: deterministic for a fixed blockDim
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
    return t;   // valid in thread 0
}

// D_B = 2 sum over pole groups of the partials, F_B likewise, N_B = Tr(D_B Q_B),
// E_B = Tr(F_B Q_B).  The pole kernels write the upper triangle of each partial
// (G[T,T] is complex symmetric); this deals the upper-triangle entries to the
// threads in row order (coalesced reads and writes), writes the upper triangles
// of D_B and F_B (k_mirror fills the lower ones for the consumers that read them)
// and forms the traces with the off-diagonal entries doubled (Q_B is symmetric).
// Windows with by_pole[w] != 0 (complex64 matrix-core passes) hold one float pair
// (Re w G, Re w z G) per pole instead, nslot poles in order.
__global__ void __launch_bounds__(1024) k_reduce(const WinDesc* __restrict__ wins, int ngroup,
                                                const double* __restrict__ Dpart, const double* __restrict__ Fpart,
                                                const float2* __restrict__ slots, int nslot,
                                                const int8_t* __restrict__ by_pole,
                                                const double* __restrict__ Q, double* __restrict__ D,
                                                double* __restrict__ F, double* __restrict__ NB,
                                                double* __restrict__ EB, const double* __restrict__ circ) {
    __shared__ double red[16];
    if (circ != nullptr && circ[2] != 0.0) return;   // the pass was skipped
    const int w = blockIdx.x;
    const WinDesc d = wins[w];
    const int k = d.k;
    const int64_t kk = (int64_t)k * k;
    const double* dp = Dpart + d.kk_off * ngroup;
    const double* fp = Fpart + d.kk_off * ngroup;
    if (by_pole != nullptr && by_pole[w] == 2) return;   // reduced by the window-fused pole kernel
    const bool poles = by_pole != nullptr && by_pole[w] == 1;
    const float2* gs = slots + d.kk_off * nslot;
    double* Dw = D + d.kk_off;
    double* Fw = F + d.kk_off;
    const double* Qw = Q + d.kk_off;
    double sn = 0.0, se = 0.0;
    // the upper-triangle entries in row order, dealt to the threads in turn (equal
    // work per thread, consecutive threads on consecutive columns): entry e is
    // (i, j) with i the last row whose start s(i) = i k - i (i - 1) / 2 is <= e
    const int nup = k * (k + 1) / 2;
    auto row_start = [k](int r) { return r * k - (r * (r - 1)) / 2; };
    for (int e = threadIdx.x; e < nup; e += blockDim.x) {
        const float b2 = 2.0f * (float)k + 1.0f;
        int i = (int)(0.5f * (b2 - sqrtf(fmaxf(b2 * b2 - 8.0f * (float)e, 0.0f))));
        i = min(max(i, 0), k - 1);
        while (i > 0 && row_start(i) > e) --i;
        while (i + 1 < k && row_start(i + 1) <= e) ++i;
        const int j = i + (e - row_start(i));
        {
            const int64_t q = (int64_t)i * k + j;
            double dv = 0.0, fv = 0.0;
            if (poles) {
                // eight loads in flight, summed in pole order: every load is issued
                // (index clamped to the last pole) and the values past the last
                // pole are masked, so no load sits behind a branch of its own
                for (int j0 = 0; j0 < nslot; j0 += 8) {
                    float2 g[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) g[u] = gs[(int64_t)min(j0 + u, nslot - 1) * kk + q];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const bool in = j0 + u < nslot;
                        dv += in ? (double)g[u].x : 0.0;
                        fv += in ? (double)g[u].y : 0.0;
                    }
                }
            } else {
                // four groups' loads in flight (clamped and masked as above)
                for (int g0 = 0; g0 < ngroup; g0 += 4) {
                    double a[4], c[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int64_t o = (int64_t)min(g0 + u, ngroup - 1) * kk + q;
                        a[u] = dp[o];
                        c[u] = fp[o];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const bool in = g0 + u < ngroup;
                        dv += in ? a[u] : 0.0;
                        fv += in ? c[u] : 0.0;
                    }
                }
            }
            dv *= 2.0;
            fv *= 2.0;
            Dw[q] = dv;
            Fw[q] = fv;
            double qv = Qw[q];
            if (j != i) qv *= 2.0;
            sn = fma(dv, qv, sn);
            se = fma(fv, qv, se);
        }
    }
    const double tn = block_sum(sn, red);
    const double te = block_sum(se, red);
    if (threadIdx.x == 0) {
        if (NB) NB[w] = tn;
        if (EB) EB[w] = te;
    }
}

// Lower triangle of D_B (and F_B if given) from the upper one, per window.
__global__ void __launch_bounds__(256) k_mirror(const WinDesc* __restrict__ wins, double* __restrict__ D,
                                                double* __restrict__ F, const double* __restrict__ skip) {
    if (skip != nullptr && *skip != 0.0) return;
    const WinDesc d = wins[blockIdx.x];
    const int k = d.k;
    double* Dw = D + d.kk_off;
    double* Fw = F ? F + d.kk_off : nullptr;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5, nty = blockDim.x >> 5;
    // column segments of the lower triangle: rows j > i read from row i (coalesced
    // writes along each lower row)
    for (int j = ty; j < k; j += nty)
        for (int i = tx; i < j; i += 32) {
            Dw[(int64_t)j * k + i] = Dw[(int64_t)i * k + j];
            if (Fw) Fw[(int64_t)j * k + i] = Fw[(int64_t)i * k + j];
        }
}

// n(g) = x_g^T D_B x_g on the owned points (src/lrft_cpp.cpp:183-191 in D form).
// One wave per owned point: lane a holds entry a of the point's row of X
// (restricted to T), y_a = sum_b D[b][a] x_b (D_B is symmetric, so every lane
// reads the same row b: contiguous addresses) with b, x_b broadcast by
// readlane, then n = sum_a x_a y_a by a fixed-shape wave reduction (bitwise
// reproducible).  Each wave takes chunks of 63 consecutive owned points: the
// chunk's row offsets and grid indices come in one load each, and the next
// point's row of X is loaded while the current one is summed, so a point waits
// on no global load of its own.  D_B is staged in LDS when k^2 doubles fit
// 56 KB, else read from global memory; rows above 64 entries take a pair loop.
constexpr int kDensWaves = 16;

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int l) {
    const int lo = __builtin_amdgcn_readlane((int)v, l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return ((int64_t)hi << 32) | (int64_t)(unsigned)lo;
}

// STAGE: D_B read from its LDS copy (a separate instantiation, so that the reads
// are LDS instructions rather than generic ones)
template <bool STAGE>
__global__ void __launch_bounds__(64 * kDensWaves) k_density(const WinDesc* __restrict__ wins,
                                                             const double* __restrict__ D,
                                                             const int64_t* __restrict__ own_g,
                                                             const int64_t* __restrict__ xptr,
                                                             const int32_t* __restrict__ xidx,
                                                             const double* __restrict__ xval,
                                                             double* __restrict__ dens,
                                                             const double* __restrict__ skip) {
    extern __shared__ double Dl[];
    if (skip != nullptr && *skip != 0.0) return;
    const int w = blockIdx.x;
    const WinDesc d = wins[w];
    const int k = d.k;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const double* Dw = D + d.kk_off;
    if constexpr (STAGE) {
        for (int i = wave; i < k; i += nwave)
            for (int j = lane; j < k; j += 64) Dl[i * k + j] = Dw[(int64_t)i * k + j];
        __syncthreads();
    }
    auto Dr = [&](int a, int b) __attribute__((always_inline)) {
        if constexpr (STAGE)
            return Dl[a * k + b];
        else
            return Dw[(int64_t)a * k + b];
    };
    for (int c0 = 63 * wave; c0 < d.nown; c0 += 63 * nwave) {
        const int cn = min(63, d.nown - c0);   // lanes 0..cn hold the cn + 1 row offsets
        const int64_t o0 = d.own_off + c0;
        // lane t: start of point t's row (lane cn: the end of the last row), grid index of point t
        const int64_t pl = lane <= cn ? xptr[o0 + lane] : 0;
        const int64_t gl = lane < cn ? own_g[o0 + lane] : 0;
        const int64_t pn = __shfl_down(pl, 1, 64);
        if (__ballot(lane < cn && pn - pl > 64) != 0) {
            // a row above 64 entries in this chunk: pair loop, point by point
            for (int t = 0; t < cn; ++t) {
                const int64_t p0 = readlane_i64(pl, t);
                const int nnz = (int)(readlane_i64(pl, t + 1) - p0);
                double acc = 0.0;
                for (int q = lane; q < nnz * nnz; q += 64) {
                    const int ai = q / nnz, bi = q - (q / nnz) * nnz;
                    const int ja = xidx[p0 + ai], jb = xidx[p0 + bi];
                    acc = fma(xval[p0 + ai] * Dr(ja, jb), xval[p0 + bi], acc);
                }
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
                const int64_t g = readlane_i64(gl, t);
                if (lane == 0) dens[g] = acc;
            }
            continue;
        }
        const int64_t pbeg = readlane_i64(pl, 0), pend = readlane_i64(pl, cn);
        if (pend == pbeg) {   // no entries in the whole chunk
            if (lane < cn) dens[gl] = 0.0;
            continue;
        }
        // every load below is issued unconditionally at an index clamped into the
        // chunk's entries [pbeg, pend) and masked afterwards, and the densities are
        // stored once per chunk: the loop then holds exactly two loads per pass,
        // so the wait before a point's sum covers its own row and not the next
        const int64_t plast = pend - 1;
        int nnz = (int)(readlane_i64(pl, 1) - pbeg);
        int ia = xidx[min(pbeg + lane, plast)];
        double va = xval[min(pbeg + lane, plast)];
        ia = lane < nnz ? ia : 0;
        va = lane < nnz ? va : 0.0;
        double res = 0.0;   // lane t: density of point t
        for (int t = 0; t < cn; ++t) {
            // next point's row, in flight while this one is summed (after the last
            // point its length reads as 0)
            const int t1 = min(t + 1, cn), t2 = min(t + 2, cn);
            const int64_t p0_n = readlane_i64(pl, t1);
            const int nnz_n = (int)(readlane_i64(pl, t2) - p0_n);
            int ia_n = xidx[min(p0_n + lane, plast)];
            double va_n = xval[min(p0_n + lane, plast)];
            // four partial sums keep four reads in flight; lanes past nnz hold
            // index 0 and value 0, so the rounded-up tail adds nothing (b + 3 < 64)
            double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
            for (int b = 0; b < nnz; b += 4) {
                const int i0 = __builtin_amdgcn_readlane(ia, b), i1 = __builtin_amdgcn_readlane(ia, b + 1);
                const int i2 = __builtin_amdgcn_readlane(ia, b + 2), i3 = __builtin_amdgcn_readlane(ia, b + 3);
                y0 = fma(Dr(i0, ia), readlane_f64(va, b), y0);
                y1 = fma(Dr(i1, ia), readlane_f64(va, b + 1), y1);
                y2 = fma(Dr(i2, ia), readlane_f64(va, b + 2), y2);
                y3 = fma(Dr(i3, ia), readlane_f64(va, b + 3), y3);
            }
            double acc = va * ((y0 + y1) + (y2 + y3));
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
            const double tot = readlane_f64(acc, 0);
            res = lane == t ? tot : res;
            nnz = nnz_n;
            ia = lane < nnz_n ? ia_n : 0;
            va = lane < nnz_n ? va_n : 0.0;
        }
        if (lane < cn) dens[gl] = res;
    }
}

// n(g) = x_g^T D_B x_g when every owned row of X is made of whole atom blocks: a
// multiple of 6 entries, each 6 in consecutive T-local columns (the 6 functions of
// one atom share its support), blocks in ascending column order.  The rows come in
// 64-point chunks, block-interleaved: block slot A of a chunk holds the 64 points'
// column indices, then their 6 x 64 values, so every load of a wave is 64 consecutive
// words (one pass over the rows, whole cache lines); a chunk's rows are padded to its
// longest with zero blocks.  One wave per chunk, one lane per point: the sum of
// x_A^T D[A, B] x_B over the blocks A <= B of its row (pairs A < B doubled: D_B is
// symmetric), D_B's 6 x 6 blocks read from LDS (or global memory when D_B does not
// fit), only entries on or above the diagonal; each lane adds in a fixed order, and
// the padding adds exact zeros after a row's own terms (bitwise reproducible).
constexpr int kDensBlkThreads = 512;

template <bool STAGE>
__global__ void __launch_bounds__(kDensBlkThreads) k_density_chunk(const WinDesc* __restrict__ wins,
                                                                   const double* __restrict__ D,
                                                                   const int64_t* __restrict__ own_g,
                                                                   const int64_t* __restrict__ chk_off,
                                                                   const int64_t* __restrict__ chk_bs,
                                                                   const int32_t* __restrict__ chk_idx,
                                                                   const double* __restrict__ chk_val,
                                                                   double* __restrict__ dens,
                                                                   const double* __restrict__ skip) {
    extern __shared__ double Dl[];
    if (skip != nullptr && *skip != 0.0) return;
    const int w = blockIdx.x;
    const WinDesc d = wins[w];
    const int k = d.k;
    const double* Dw = D + d.kk_off;
    if constexpr (STAGE) {
        // eight loads in flight per thread (clamped index, masked store): a plain
        // copy loop waits for each load before the next
        const int kk = k * k, nt = blockDim.x;
        for (int e0 = threadIdx.x; e0 < kk; e0 += 8 * nt) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = Dw[min(e0 + u * nt, kk - 1)];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + u * nt < kk) Dl[e0 + u * nt] = v[u];
        }
        __syncthreads();
    }
    auto Dr = [&](int a, int b) __attribute__((always_inline)) {
        if constexpr (STAGE)
            return Dl[a * k + b];
        else
            return Dw[(int64_t)a * k + b];
    };
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), nwave = blockDim.x >> 6;
    const int64_t c0 = chk_off[w], c1 = chk_off[w + 1];
    for (int64_t c = c0 + wave; c < c1; c += nwave) {
        const int64_t b0 = chk_bs[c];
        const int nb = (int)(chk_bs[c + 1] - b0);
        const int p = (int)(c - c0) * 64 + lane;
        double acc = 0.0;
        if (nb > 0) {
            // block A of this lane's row: column xi[64 A], values xv[64 (6 A + i)]; each
            // block is loaded one step ahead of its use (the next block clamped to the
            // last), so a load is in flight while the previous block's terms are summed
            const int32_t* xi = chk_idx + b0 * 64 + lane;
            const double* xv = chk_val + b0 * 6 * 64 + lane;
            int sa = xi[0];
            double xa[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) xa[i] = xv[64 * i];
#pragma unroll 1
            for (int A = 0; A < nb; ++A) {
                const int An = min(A + 1, nb - 1);
                const int sn = xi[64 * An];
                double xn[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) xn[i] = xv[64 * (6 * An + i)];
                // diagonal block, every entry read from the upper triangle
                double t = 0.0;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    double y = 0.0;
#pragma unroll
                    for (int j = 0; j < 6; ++j) y = fma(Dr(sa + (i < j ? i : j), sa + (i < j ? j : i)), xa[j], y);
                    t = fma(xa[i], y, t);
                }
                acc += t;
                int sb = sn;
                double xb[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) xb[j] = xn[j];
#pragma unroll 1
                for (int B = A + 1; B < nb; ++B) {
                    const int Bn = min(B + 1, nb - 1);
                    const int sbn = xi[64 * Bn];
                    double xbn[6];
#pragma unroll
                    for (int j = 0; j < 6; ++j) xbn[j] = xv[64 * (6 * Bn + j)];
                    // a padding block (zero values, column 0) after a real one reads entries
                    // below the diagonal, which passes leave as the plan created them (zero)
                    // or as the mirror kernel filled them, and multiplies them by zero
                    double u = 0.0;
#pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        double y = 0.0;
#pragma unroll
                        for (int j = 0; j < 6; ++j) y = fma(Dr(sa + i, sb + j), xb[j], y);
                        u = fma(xa[i], y, u);
                    }
                    acc = fma(2.0, u, acc);
                    sb = sbn;
#pragma unroll
                    for (int j = 0; j < 6; ++j) xb[j] = xbn[j];
                }
                sa = sn;
#pragma unroll
                for (int i = 0; i < 6; ++i) xa[i] = xn[i];
            }
        }
        if (p < d.nown) dens[own_g[d.own_off + p]] = acc;
    }
}

// ---------------------------------------------------------------------------
// SCF density mixing (the caller's DIIS step): out = max(sum_h c_h (N_h + mix R_h), lo),
// then out *= ntot / (sum(out) dV).  Fixed grid and block shapes give the sum a fixed
// order (bitwise reproducible, also when the launches are replayed from a graph), and
// every partial slot is written, so nothing has to be cleared beforehand.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMixThreads) k_mix(const double* __restrict__ coef, int nh,
                                                     const double* __restrict__ Nh, const double* __restrict__ Rh,
                                                     int64_t ng, double mix, double lo, double* __restrict__ out,
                                                     double* __restrict__ part) {
    __shared__ double red[16];
    double s = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * kMixThreads + threadIdx.x; g < ng; g += (int64_t)kMixBlocks * kMixThreads) {
        double v = 0.0;
        for (int h = 0; h < nh; ++h) v = fma(coef[h], fma(mix, Rh[(int64_t)h * ng + g], Nh[(int64_t)h * ng + g]), v);
        v = fmax(v, lo);
        out[g] = v;
        s += v;
    }
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// every block forms the same total from the kMixBlocks partials in the same order
__global__ void __launch_bounds__(kMixThreads) k_mix_scale(const double* __restrict__ part, double ntot, double dV,
                                                           int64_t ng, double* __restrict__ out) {
    __shared__ double red[16];
    __shared__ double scale;
    double s = 0.0;
    for (int b = threadIdx.x; b < kMixBlocks; b += kMixThreads) s += part[b];
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) scale = ntot / (t * dV);
    __syncthreads();
    const double sc = scale;
    for (int64_t g = (int64_t)blockIdx.x * kMixThreads + threadIdx.x; g < ng; g += (int64_t)kMixBlocks * kMixThreads)
        out[g] *= sc;
}

// ---------------------------------------------------------------------------
// SCF grid work around the density step (lcao3d.py:244-281, three_d.py:132-148):
// the LDA potential, the residual with its DIIS dot products and the
// Harris-Foulkes integrals, each one pass over the grid.  Sums go through
// kMixBlocks fixed partials per quantity and one finishing block (fixed order).
// ---------------------------------------------------------------------------
struct LdaXc {
    double exc, vxc;
};

// Slater exchange + Perdew-Zunger 1981 correlation of one density value, with
// the clamp at 1e-14 (three_d.py:132-148)
__device__ __forceinline__ LdaXc lda_pz81(double n) {
    constexpr double kPi = 3.14159265358979323846;
    constexpr double gam = -0.1423, b1 = 1.0529, b2 = 0.3334;
    constexpr double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    n = fmax(n, 1e-14);
    const double n13 = cbrt(n);
    const double rs = cbrt(3.0 / (4.0 * kPi)) / n13;
    const double ex = -0.75 * cbrt(3.0 / kPi) * n13;
    const double vx = (4.0 / 3.0) * ex;
    double ec, vc;
    if (rs >= 1.0) {
        const double srs = sqrt(rs);
        const double den = 1.0 + b1 * srs + b2 * rs;
        ec = gam / den;
        vc = ec * (1.0 + (7.0 / 6.0) * b1 * srs + (4.0 / 3.0) * b2 * rs) / den;
    } else {
        const double lnr = log(rs);
        ec = A * lnr + B + C * rs * lnr + D * rs;
        vc = A * lnr + (B - A / 3.0) + (2.0 / 3.0) * C * rs * lnr + (2.0 * D - C) / 3.0 * rs;
    }
    return {ex + ec, vx + vc};
}

// v = v_ion + v_H + v_xc(n)
__global__ void __launch_bounds__(kMixThreads) k_lda_potential(const double* __restrict__ n,
                                                               const double* __restrict__ vion,
                                                               const double* __restrict__ vh, int64_t ng,
                                                               double* __restrict__ v) {
    for (int64_t g = (int64_t)blockIdx.x * kMixThreads + threadIdx.x; g < ng; g += (int64_t)kMixBlocks * kMixThreads)
        v[g] = vion[g] + vh[g] + lda_pz81(n[g]).vxc;
}

// r = scale nout - nin; Rh[slot] = r, Nh[slot] = nin; partials of sum |r| (row 0)
// and of R_q . r for every history row q < nh (rows 1..nh; row slot is r . r)
// block_sum of each of the nq per-thread values v[q] into part[q kMixBlocks + block]: the
// same additions (each wave's shuffle tree, then the wave partials in order from zero), the
// rows' trees back to back and one barrier for all of them
template <int N>
__device__ __forceinline__ void rows_to_parts(const double (&v)[N], int nq, double (&red)[kMaxHist + 1][kMixThreads / 64],
                                              double* __restrict__ part) {
    static_assert(N <= kMaxHist + 1, "rows");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        if (q >= nq) break;
        double x = v[q];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[q][wave] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < nq) {
        double t = 0.0;
        for (int w = 0; w < kMixThreads / 64; ++w) t += red[threadIdx.x][w];
        part[(int64_t)threadIdx.x * kMixBlocks + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kMixThreads) k_scf_residual(const double* __restrict__ nout, double scale,
                                                              const double* __restrict__ nin, int64_t ng, int nh,
                                                              int slot, double* __restrict__ Nh,
                                                              double* __restrict__ Rh, double* __restrict__ part) {
    __shared__ double red2[kMaxHist + 1][kMixThreads / 64];
    double s[kMaxHist + 1];
#pragma unroll
    for (int q = 0; q <= kMaxHist; ++q) s[q] = 0.0;
    double* Rs = Rh + (int64_t)slot * ng;
    for (int64_t g = (int64_t)blockIdx.x * kMixThreads + threadIdx.x; g < ng; g += (int64_t)kMixBlocks * kMixThreads) {
        const double x = nin[g];
        const double r = scale * nout[g] - x;
        Rs[g] = r;
        Nh[(int64_t)slot * ng + g] = x;
        s[0] += fabs(r);
#pragma unroll
        for (int q = 0; q < kMaxHist; ++q)
            if (q < nh) s[1 + q] = fma(q == slot ? r : Rh[(int64_t)q * ng + g], r, s[1 + q]);
    }
    rows_to_parts(s, nh + 1, red2, part);
}

// partials of sum n v_H (row 0), sum n e_xc(n) (row 1), sum n (v_H + v_xc(n)) (row 2)
__global__ void __launch_bounds__(kMixThreads) k_harris(const double* __restrict__ n, const double* __restrict__ vh,
                                                        int64_t ng, double* __restrict__ part) {
    __shared__ double red2[kMaxHist + 1][kMixThreads / 64];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * kMixThreads + threadIdx.x; g < ng; g += (int64_t)kMixBlocks * kMixThreads) {
        const double x = n[g], h = vh[g];
        const LdaXc l = lda_pz81(x);
        s0 = fma(x, h, s0);
        s1 = fma(x, l.exc, s1);
        s2 = fma(h + l.vxc, x, s2);
    }
    const double sv[3] = {s0, s1, s2};
    rows_to_parts(sv, 3, red2, part);
}

// out[q] = sum_b part[q kMixBlocks + b] for q < nq (one block, fixed order)
__global__ void __launch_bounds__(kMixThreads) k_rows_total(const double* __restrict__ part, int nq,
                                                            double* __restrict__ out) {
    // block_sum's additions for every row (each thread's strided sum, each wave's shuffle
    // tree, the wave partials in order), the rows' loads and trees back to back with one
    // barrier for all of them
    constexpr int kWaves = kMixThreads / 64;
    __shared__ double red[kMaxHist + 1][kWaves];   // nq <= kMaxHist + 1 (the entry points check it)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < nq; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < kMixBlocks; b += kMixThreads) s += part[(int64_t)q * kMixBlocks + b];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) red[q][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x < nq) {
        double t = 0.0;
        for (int w = 0; w < kWaves; ++w) t += red[threadIdx.x][w];
        out[threadIdx.x] = t;
        __threadfence_system();   // out may be host memory read after the stream's wait
    }
}

// sum of nwin values in a fixed order into out[0] (one block)
__global__ void k_sum2(const double* __restrict__ a, const double* __restrict__ b, int n, double* __restrict__ out,
                       const double* __restrict__ skip) {
    __shared__ double red[16];
    if (skip != nullptr && *skip != 0.0) return;
    double sa = 0.0, sb = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        sa += a[q];
        sb += b[q];
    }
    const double ta = block_sum(sa, red);
    const double tb = block_sum(sb, red);
    if (threadIdx.x == 0) {
        out[0] = ta;
        out[1] = tb;
        __threadfence_system();   // out may be mapped host memory
    }
}

// ---------------------------------------------------------------------------
// The circle of a pass in device memory and the warm mu search between the
// queued passes of a device step: one thread runs the search of lres_musearch.h
// (the host loop's code) on the sums the pass left in device memory and writes the
// next pass's circle, or marks the remaining queued passes as skipped.
// ---------------------------------------------------------------------------
__global__ void k_set_circle(double* __restrict__ circ, double c, double R, double skip, double skip_density) {
    circ[0] = c;
    circ[1] = R;
    circ[2] = skip;
    circ[3] = skip_density;
}

// a trial pass at mu (the same c and R as the host's run_pass); trials form no density
__device__ void circle_at(double* circ, double mu, double e_lo) {
    circ[0] = 0.5 * (mu + e_lo);
    circ[1] = 0.5 * (mu - e_lo);
    circ[2] = 0.0;
    circ[3] = 1.0;
}

__device__ void circle_off(double* circ) {
    circ[2] = 1.0;
    circ[3] = 1.0;
}

// ---------------------------------------------------------------------------
// Keep-edge bound.  With S = X^T X dV and H(v) = T + X^T diag(v) X dV, a potential change
// dv moves every window pencil's eigenvalues by at most max_g |dv(g)| (for any c,
// |c^T X^T diag(dv) X c dV| <= max |dv| c^T S c; min-max).  So with the lowest window
// eigenvalue bracketed as [lo, hi] at a reference potential v_ref, the keep-edge test's
// outcome at v (no eigenvalue below e_lo + m / 3, one below e_lo + 2 m) is proved when
// lo - delta > e_lo + m / 3 and hi + delta < e_lo + 2 m, delta = max |v - v_ref|, and the
// two inertia counts are then not needed.  edge = {lo, hi, delta, tests skipped, tests
// checked}.

// vcur = v and per-block maxima of |v - vref| (a NaN anywhere propagates)
// The largest of a and b, NaN when either is NaN (so that a NaN anywhere in the
// potential reaches the bound, which then proves nothing)
__device__ __forceinline__ double edge_max(double a, double b) {
    return (a != a || b != b) ? a + b : (a > b ? a : b);
}

// The potential v copied to vcur, and per block the largest rise max(v - vref) and the
// largest fall max(vref - v) (part[2 b], part[2 b + 1]).  With H_W(v) = T_W + X^T diag(v)
// X dV and S_W = X^T X dV, x^T dH x / x^T S x is a weighted mean of v - vref, so every
// window eigenvalue moves by at least -fall and at most +rise.
__global__ void __launch_bounds__(256) k_edge_delta(const double* __restrict__ v, double* __restrict__ vcur,
                                                    const double* __restrict__ vref, int64_t ng,
                                                    double* __restrict__ part) {
    __shared__ double red[2][256];
    double up = -INFINITY, dn = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < ng; i += (int64_t)gridDim.x * 256) {
        const double x = v[i];
        vcur[i] = x;
        const double d = x - vref[i];
        up = edge_max(up, d);
        dn = edge_max(dn, -d);
    }
    red[0][threadIdx.x] = up;
    red[1][threadIdx.x] = dn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] = edge_max(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = edge_max(red[1][threadIdx.x], red[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0];
        part[2 * blockIdx.x + 1] = red[1][0];
    }
}

// the blocks' rise and fall into edge[2] and edge[5]
__global__ void __launch_bounds__(256) k_edge_final(const double* __restrict__ part, int nb,
                                                    double* __restrict__ edge) {
    __shared__ double red[2][256];
    double up = -INFINITY, dn = -INFINITY;
    for (int i = threadIdx.x; i < nb; i += 256) {
        up = edge_max(up, part[2 * i]);
        dn = edge_max(dn, part[2 * i + 1]);
    }
    red[0][threadIdx.x] = up;
    red[1][threadIdx.x] = dn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] = edge_max(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = edge_max(red[1][threadIdx.x], red[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        edge[2] = red[0][0];
        edge[5] = red[1][0];
    }
}

// a new reference: the lowest eigenvalue bracket [lo, hi] at the current potential
__global__ void k_edge_ref(double* __restrict__ edge, double lo, double hi) {
    edge[0] = lo;
    edge[1] = hi;
    edge[2] = 0.0;
    edge[5] = 0.0;
}

// the keep-edge test at e_lo decided from the bound when it can be: then the inertia
// totals it would give (0 below e_lo + m / 3, 1 below e_lo + 2 m) go to itot and *skip
// tells the inertia kernels to leave at once; safety covers the rounding of the
// factorisations the bound stands in for
// The keep-edge test of a device step, set up on the device from the parameter block:
// its shifts e_lo + m / 3 and e_lo + 2 m into sig (as the host forms them), the totals
// zeroed (no memset: the step may be captured into a graph), and with edge given the
// decision from the bound where it proves it (a reference bracket of NaN proves nothing)
__global__ void k_step_edge(const double* __restrict__ prm, double* __restrict__ sig,
                            unsigned long long* __restrict__ itot, double* __restrict__ edge, double safety,
                            int* __restrict__ skip) {
#pragma clang fp contract(off)
    const double e_lo = prm[kStepPrmElo], margin = prm[kStepPrmMargin];
    sig[0] = e_lo + margin / 3;
    sig[1] = e_lo + 2 * margin;
    itot[0] = 0;
    itot[1] = 0;
    bool proved = false;
    if (edge != nullptr) {
        // the smallest eigenvalue falls by at most edge[5] and rises by at most edge[2]
        const double lo = edge[0] - edge[5], hi = edge[1] + edge[2];
        proved = lo > e_lo + margin / 3 + safety && hi < e_lo + 2 * margin - safety;
        if (proved) {
            itot[1] = 1;
            edge[3] += 1.0;
        }
        edge[4] += 1.0;
    }
    *skip = proved ? 1 : 0;
}

// ---------------------------------------------------------------------------
// lambda_min on the device: the host search of lres_lambda_min (Rayleigh bound, a
// bracket from counts at hi - 1, hi - 3, hi - 7, ..., then multisection rounds over the
// windows that can still hold the minimum) with every decision taken by one-block
// controller kernels between the count launches, so the whole search waits once.  The
// arithmetic is the host's operation for operation (no contraction), so both give the
// same bracket bit for bit.

// shifts of a multisection round over ncand windows: as lres_lambda_min's
__device__ int lmin_shifts(int ncand) {
    const int k = kLminWaves / (ncand > 1 ? ncand : 1);
    return k < kLminMinK ? kLminMinK : (k > kLminMaxK ? kLminMaxK : k);
}

// the next round's K shifts over [lo, hi] (totals zeroed), by the whole block; s->lo,
// s->hi and s->ncand are set and visible to every thread
__device__ void lmin_setup_round(LminState* s, double* sig, unsigned long long* itot) {
#pragma clang fp contract(off)
    const int K = lmin_shifts(s->ncand);
    const double lo = s->lo, hi = s->hi;
    for (int q = threadIdx.x; q < K; q += blockDim.x) {
        sig[q] = lo + (hi - lo) * (q + 1) / (K + 1);
        itot[q] = 0;
    }
    if (threadIdx.x == 0) s->K = K;
}

// the Rayleigh bound hi0 = min_w min_a H_aa / S_aa (+ 1e-12 (1 + |hi0|)) and the probe
// shifts hi0, hi0 - 1, hi0 - 3, hi0 - 7, ... (kLminProbe below hi0), totals zeroed
__global__ void __launch_bounds__(256) k_lmin_probe(const double* __restrict__ rq, int nwin, LminState* __restrict__ s,
                                                    double tol, double* __restrict__ sig,
                                                    unsigned long long* __restrict__ itot) {
#pragma clang fp contract(off)
    __shared__ double red[256];
    double m = 1.0 / 0.0;
    for (int w = threadIdx.x; w < nwin; w += 256) m = fmin(m, rq[w]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int q = 128; q > 0; q >>= 1) {
        if (threadIdx.x < q) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + q]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double hi = red[0];
    hi += 1e-12 * (1.0 + fabs(hi));
    s->tol = tol;
    s->hi = hi;
    s->lo = hi;
    s->ncand = s->K = 0;
    s->done = s->fail = 0;
    sig[0] = hi;
    itot[0] = 0;
    double step = 1.0, lo = hi - step;
    for (int j = 0; j < kLminProbe; ++j) {
        sig[1 + j] = lo;
        itot[1 + j] = 0;
        step *= 2.0;
        lo -= step;
    }
}

// the bracket from the probe counts (the first probe counting nothing is lo, the one
// above it hi), all windows as candidates and the first round's shifts
__global__ void __launch_bounds__(1024) k_lmin_bracket(LminState* __restrict__ s, double* __restrict__ sig,
                                                       unsigned long long* __restrict__ itot,
                                                       const int32_t* __restrict__ ids, int nwin,
                                                       int32_t* __restrict__ cand) {
#pragma clang fp contract(off)
    for (int w = threadIdx.x; w < nwin; w += 1024) cand[w] = ids[w];
    if (threadIdx.x == 0) {
        if (itot[0] == 0) {            // no eigenvalue below the Rayleigh bound: S_W not definite
            s->fail = 1;
            s->done = 1;
        } else {
            int j = 0;
            while (j < kLminProbe && itot[1 + j] != 0) ++j;
            if (j == kLminProbe) {     // the spectrum reaches further down: the host search
                s->fail = 2;
                s->done = 1;
            } else {
                s->lo = sig[1 + j];
                s->hi = sig[j];
                s->ncand = nwin;
                s->done = s->hi - s->lo > s->tol ? 0 : 1;
            }
        }
    }
    __syncthreads();
    if (!s->done) lmin_setup_round(s, sig, itot);
}

// one multisection round's outcome: the new bracket, the candidates that count an
// eigenvalue below its upper end (while more than 64 take part), the next round's shifts
__global__ void __launch_bounds__(1024) k_lmin_round(LminState* __restrict__ s, double* __restrict__ sig,
                                                     unsigned long long* __restrict__ itot,
                                                     const int32_t* __restrict__ counts, int kcap,
                                                     int32_t* __restrict__ cand, int32_t* __restrict__ tmp) {
#pragma clang fp contract(off)
    __shared__ int first_s, base_s, next_s;
    __shared__ int scan[1024];
    if (s->done) return;
    const int ncand = s->ncand, K = s->K;
    if (threadIdx.x == 0) {
        first_s = K;
        base_s = 0;
    }
    __syncthreads();
    // the first shift with a positive total
    for (int q = threadIdx.x; q < K; q += 1024)
        if (itot[q] != 0) atomicMin(&first_s, q);
    __syncthreads();
    const int fp = first_s < K ? first_s : -1;
    if (fp >= 0 && ncand > 64) {
        // ordered compaction of the candidates with a positive count at shift fp
        for (int c0 = 0; c0 < ncand; c0 += 1024) {
            const int c = c0 + threadIdx.x;
            const int keep = c < ncand && counts[(int64_t)cand[c] * kcap + fp] > 0 ? 1 : 0;
            scan[threadIdx.x] = keep;
            __syncthreads();
            for (int o = 1; o < 1024; o <<= 1) {
                const int x = threadIdx.x >= o ? scan[threadIdx.x - o] : 0;
                __syncthreads();
                scan[threadIdx.x] += x;
                __syncthreads();
            }
            if (keep) tmp[base_s + scan[threadIdx.x] - 1] = cand[c];
            __syncthreads();
            if (threadIdx.x == 0) base_s += scan[1023];
            __syncthreads();
        }
        const int nn = base_s;
        if (nn > 0) {
            for (int c = threadIdx.x; c < nn; c += 1024) cand[c] = tmp[c];
            if (threadIdx.x == 0) s->ncand = nn;
        }
    }
    if (threadIdx.x == 0) {
        // as the host's scan: shifts counting nothing raise lo, the first positive one is hi
        const double lo = s->lo, hi = s->hi;
        const double nlo = fp < 0 ? sig[K - 1] : (fp > 0 ? sig[fp - 1] : lo);
        const double nhi = fp < 0 ? hi : sig[fp];
        if (nlo == lo && nhi == hi) {   // no progress at double resolution
            s->done = 1;
        } else {
            s->lo = nlo;
            s->hi = nhi;
            s->done = nhi - nlo > s->tol ? 0 : 1;
        }
        next_s = s->done ? 0 : 1;
    }
    __syncthreads();
    if (next_s) lmin_setup_round(s, sig, itot);
}

__global__ void k_lmin_snap(LminState* __restrict__ s, double width, double* __restrict__ sig,
                            unsigned long long* __restrict__ itot, int* __restrict__ skip) {
    if (!s->done || s->fail) {   // not finished in this batch (or failed): nothing to snap
        *skip = 1;
        return;
    }
    double cell = 0.0;
    const double g = lmin_grid(s->lo, s->hi, width, &cell);
    if (isnan(g)) {
        s->cell = cell;
        *skip = 1;
    } else {
        sig[0] = g;
        itot[0] = 0;
        *skip = 0;
    }
}

__global__ void k_lmin_snap_end(LminState* __restrict__ s, double width, const double* __restrict__ sig,
                                const unsigned long long* __restrict__ itot, const int* __restrict__ skip) {
    if (*skip != 0) return;
    s->cell = itot[0] > 0 ? sig[0] - width : sig[0];   // an eigenvalue below g: the cell under it
}

// itot (may be null): inertia totals at (e_lo + margin / 3, e_lo + 2 margin); the
// filter edge is kept only when the first is zero and the second positive
// (lcao3d.py:420), else the step stops with kWarmEdge before any pass
// (the step's inputs from the parameter block prm, kStepPrm* order, written by the host
// before the step is launched or its graph replayed)
__global__ void k_mu_begin(StepState* __restrict__ st, const unsigned long long* __restrict__ itot,
                           const double* __restrict__ prm, double* __restrict__ circ) {
    WarmSearch& s = st->s;
    warm_begin_keep(s, prm[kStepPrmN], prm[kStepPrmElo], prm[kStepPrmMu0], prm[kStepPrmSlope], prm[kStepPrmKeepTol],
                    prm[kStepPrmFtol], prm[kStepPrmXtol], (int)prm[kStepPrmLag]);
    st->final_pass = 0;
    st->final_ran = 0;
    if (itot != nullptr && !(itot[0] == 0 && itot[1] > 0)) s.phase = kWarmEdge;
    if (warm_terminal(s))
        circle_off(circ);
    else
        circle_at(circ, s.x, s.e_lo);
}

__device__ void step_final(StepState* st, double* circ) {
    WarmSearch& s = st->s;
    if (s.phase == kWarmDone) {
        if (s.mu == s.last_mu) {
            circ[2] = 1.0;     // the density from the last trial's D_B
        } else {
            circle_at(circ, s.mu, s.e_lo);
            st->final_pass = 1;
        }
        circ[3] = 0.0;
    } else {
        circle_off(circ);      // still searching (resumed by the next batch) or failed
    }
}

// one controller stage on the sums (sn, se) of the slot's pass (valid when circ[2] was 0
// before it ran: `ran`)
__device__ void step_stage(StepState* st, int stage, bool ran, double sn, double se, double* circ,
                           unsigned long long* pin) {
    WarmSearch& s = st->s;
    if (stage == kStepFeed || stage == kStepFeedFinal) {
        if (ran) {
            warm_feed(s, sn, se);
            if (warm_terminal(s))
                circle_off(circ);
            else
                circle_at(circ, s.x, s.e_lo);
        }
        if (stage == kStepFeedFinal) step_final(st, circ);
    } else if (stage == kStepPublish) {
        if (st->final_pass && ran) {
            s.last_mu = s.mu;
            s.last_nsum = sn;
            s.last_esum = se;
            s.last_f = 2.0 * sn - s.N;
            st->final_pass = 0;
            st->final_ran += 1;
        }
        // the state to mapped host memory, read after the stream's wait
        const unsigned long long* src = reinterpret_cast<const unsigned long long*>(st);
        for (int i = 0; i < (int)(sizeof(StepState) / 8); ++i) pin[i] = src[i];
        __threadfence_system();
    } else if (stage == kStepResume) {
        if (warm_terminal(s))
            circle_off(circ);
        else
            circle_at(circ, s.x, s.e_lo);
    }
}

__global__ void k_mu_step(StepState* __restrict__ st, int stage, double* __restrict__ circ,
                          unsigned long long* __restrict__ pin) {
    step_stage(st, stage, false, 0.0, 0.0, circ, pin);
}

// The sums of N_B and E_B of a device step's slot (the same fixed-order sum as k_sum2)
// followed by the controller stage in thread 0; every thread reads circ[2] before
// thread 0 rewrites it.
__global__ void k_sum_step(const double* __restrict__ a, const double* __restrict__ b, int n,
                           double* __restrict__ out, StepState* __restrict__ st, int stage, double* __restrict__ circ,
                           unsigned long long* __restrict__ pin) {
    __shared__ double red[16];
    const bool ran = circ[2] == 0.0;
    double ta = 0.0, tb = 0.0;
    if (ran) {
        double sa = 0.0, sb = 0.0;
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            sa += a[q];
            sb += b[q];
        }
        ta = block_sum(sa, red);
        tb = block_sum(sb, red);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (ran) {
            out[0] = ta;
            out[1] = tb;
        }
        step_stage(st, stage, ran, ta, tb, circ, pin);
    }
}

// ---------------------------------------------------------------------------
// inertia: unpivoted LDL^T of sigma S_W - H_W, count of positive pivots
// = number of window eigenvalues below sigma (Sylvester)
// ---------------------------------------------------------------------------
template <int T>
__global__ void __launch_bounds__(1024) k_inertia(const WinDesc* __restrict__ wins, const int32_t* __restrict__ ids,
                                                  const double* __restrict__ SW, const double* __restrict__ HW,
                                                  int G, const double* __restrict__ sigma, int nshift,
                                                  int32_t* __restrict__ counts, unsigned long long* __restrict__ total,
                                                  const int* __restrict__ skip) {
    __shared__ double rowbuf[2][kMaxG * T];
    if (skip != nullptr && *skip != 0) return;   // decided by the keep-edge bound
    __shared__ double colbuf[2][kMaxG * T];
    const int w = ids[blockIdx.x];
    const int s_id = blockIdx.y;
    const WinDesc d = wins[w];
    const int n = d.n;
    const int ty = threadIdx.x / G, tx = threadIdx.x - ty * G;
    const double sg = sigma[s_id];
    const double* S = SW + d.mat_off;
    const double* H = HW + d.mat_off;
    double A[T][T];
#pragma unroll
    for (int i = 0; i < T; ++i) {
        const int r = ty + G * i;
#pragma unroll
        for (int j = 0; j < T; ++j) {
            const int col = tx + G * j;
            A[i][j] = (r < n && col < n) ? sg * S[(int64_t)r * n + col] - H[(int64_t)r * n + col]
                                         : (r == col ? 1.0 : 0.0);
        }
    }
    int npos = 0;
#pragma unroll
    for (int PI = 0; PI < T; ++PI) {
        const int pr_end = min(G, n - G * PI);
        for (int pr = 0; pr < pr_end; ++pr) {
            const int p = G * PI + pr, buf = p & 1;
            if (ty == pr) {
#pragma unroll
                for (int j = 0; j < T; ++j) rowbuf[buf][tx + G * j] = A[PI][j];
            }
            if (tx == pr) {
#pragma unroll
                for (int i = 0; i < T; ++i) colbuf[buf][ty + G * i] = A[i][PI];
            }
            __syncthreads();
            const double piv = rowbuf[buf][p];
            npos += piv > 0.0 ? 1 : 0;
            const double inv = 1.0 / piv;
            double rs[T], cc[T];
#pragma unroll
            for (int j = 0; j < T; ++j) rs[j] = rowbuf[buf][tx + G * j] * inv;
#pragma unroll
            for (int i = 0; i < T; ++i) cc[i] = colbuf[buf][ty + G * i];
#pragma unroll
            for (int i = 0; i < T; ++i)
#pragma unroll
                for (int j = 0; j < T; ++j) A[i][j] = fma(-cc[i], rs[j], A[i][j]);
        }
    }
    if (threadIdx.x == 0) {
        if (counts) counts[(int64_t)w * nshift + s_id] = npos;
        atomicAdd(&total[s_id], (unsigned long long)npos);
    }
}

// ---------------------------------------------------------------------------
// launchers (host side of this translation unit)
// ---------------------------------------------------------------------------
}  // namespace lres

using namespace lres;

hipError_t launch_diis_mix(const double* coef, int nh, const double* Nh, const double* Rh, int64_t ng, double mix,
                           double lo, double ntot, double dV, double* out, double* part, hipStream_t st) {
    hipLaunchKernelGGL(k_mix, dim3(kMixBlocks), dim3(kMixThreads), 0, st, coef, nh, Nh, Rh, ng, mix, lo, out, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mix_scale, dim3(kMixBlocks), dim3(kMixThreads), 0, st, part, ntot, dV, ng, out);
    return hipGetLastError();
}

hipError_t launch_lda_potential(const double* n, const double* vion, const double* vh, int64_t ng, double* v,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_lda_potential, dim3(kMixBlocks), dim3(kMixThreads), 0, st, n, vion, vh, ng, v);
    return hipGetLastError();
}

hipError_t launch_scf_residual(const double* nout, double scale, const double* nin, int64_t ng, int nh, int slot,
                               double* Nh, double* Rh, double* part, double* stats, hipStream_t st) {
    hipLaunchKernelGGL(k_scf_residual, dim3(kMixBlocks), dim3(kMixThreads), 0, st, nout, scale, nin, ng, nh, slot,
                       Nh, Rh, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rows_total, dim3(1), dim3(kMixThreads), 0, st, part, nh + 1, stats);
    return hipGetLastError();
}

hipError_t launch_harris(const double* n, const double* vh, int64_t ng, double* part, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_harris, dim3(kMixBlocks), dim3(kMixThreads), 0, st, n, vh, ng, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rows_total, dim3(1), dim3(kMixThreads), 0, st, part, 3, out);
    return hipGetLastError();
}

hipError_t launch_gather_csr(const WinDesc* wins, int nwin, const int64_t* Wflat, const int64_t* Woff,
                             const int32_t* pos, const int64_t* Hp, const int64_t* Hi, const double* Hv,
                             double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_csr, dim3(nwin), dim3(256), 0, st, wins, Wflat, Woff, pos, Hp, Hi, Hv, out);
    return hipGetLastError();
}

// Two register budgets per tile size T: MAXT = 1024 lets the compiler use at
// most 128 VGPRs (4 waves per SIMD, several blocks per CU), MAXT = 256 up to
// 512 (fewer, fatter waves).  The autotuner times both.
template <typename R>
static hipError_t launch_pole_t(int T, int G, int wide, const WinDesc* wins, const int32_t* ids, int nids,
                                const double* SW, const double* HW, int M, int ngroup, const double* circ,
                                double* Dpart, double* Fpart, hipStream_t st) {
    if (G < 1 || G > kMaxG) return hipErrorInvalidValue;
    if (wide && G * G > 256) return hipErrorInvalidValue;
    dim3 grid(nids, ngroup), blk(G * G);
#define LRES_POLE_CASE(TT)                                                                                    \
    case TT:                                                                                                  \
        if (wide)                                                                                             \
            hipLaunchKernelGGL((k_pole_gj<TT, 256, R>), grid, blk, 0, st, wins, ids, SW, HW, M, ngroup, G, circ,  \
                               Dpart, Fpart);                                                                 \
        else                                                                                                  \
            hipLaunchKernelGGL((k_pole_gj<TT, 1024, R>), grid, blk, 0, st, wins, ids, SW, HW, M, ngroup, G, circ, \
                               Dpart, Fpart);                                                                 \
        break;
    switch (T) {
        LRES_POLE_CASE(2)
        LRES_POLE_CASE(3)
        LRES_POLE_CASE(4)
        LRES_POLE_CASE(5)
        LRES_POLE_CASE(6)
        default: return hipErrorInvalidValue;
    }
#undef LRES_POLE_CASE
    return hipGetLastError();
}

hipError_t launch_pole_gj(int T, int G, int wide, int fp32, const WinDesc* wins, const int32_t* ids, int nids,
                          const double* SW, const double* HW, int M, int ngroup, const double* circ,
                          double* Dpart, double* Fpart, hipStream_t st) {
    if (fp32) return launch_pole_t<float>(T, G, wide, wins, ids, nids, SW, HW, M, ngroup, circ, Dpart, Fpart, st);
    return launch_pole_t<double>(T, G, wide, wins, ids, nids, SW, HW, M, ngroup, circ, Dpart, Fpart, st);
}

int mfma_acc_regs(int fp32, int nt, int nwc) {
    if (nwc == 0) nwc = 1;   // the window-fused form: one wave per matrix
    // the k-th tile column of a wave keeps the tile rows below min(nt, nwc (k + 1))
    const int tj = (nt + nwc - 1) / nwc;
    int tiles = 0;
    for (int j = 0; j < tj; ++j) tiles += nwc * (j + 1) < nt ? nwc * (j + 1) : nt;
    return tiles * (fp32 ? 8 : 16);
}

// nwc 0: the window-fused complex64 form (up to 8 tile columns)
static bool mfma_nwc_ok(int fp32, int nt, int nwc) {
    if (nwc == 0) return fp32 && nt >= 2 && nt <= 8;
    if (nt > 8) return nt <= 16 && nwc == 8;
    return nt >= 2 && nwc >= 1 && nwc <= nt;
}

hipError_t launch_pole_mfma(int fp32, int nt, int nwc, const WinDesc* wins, const int32_t* ids, int nids,
                            const double* SW, const double* HW, int M, int ngroup, const double* circ,
                            double* Dpart, double* Fpart, int acc, WinOut out, hipStream_t st) {
    if (!mfma_nwc_ok(fp32, nt, nwc)) return hipErrorInvalidValue;
#define LRES_MFMA_CALL(r, NT)                                                                                  \
    if (nt == NT && (fp32 != 0) == (sizeof(r) == 4))                                                           \
        return lres_mfma_part_##r##_##NT(nwc, wins, ids, nids, SW, HW, M, ngroup, circ, Dpart, Fpart, acc, out, \
                                         st);
    LRES_MFMA_PARTS(LRES_MFMA_CALL)
#undef LRES_MFMA_CALL
    return hipErrorInvalidValue;
}

hipError_t launch_inertia_mfma(int nt, const WinDesc* wins, const int32_t* ids, int nids, const double* SW,
                               const double* HW, const double* sigma, int nshift, int32_t* counts,
                               unsigned long long* total, const int* skip, hipStream_t st, const int* lin) {
#define LRES_INERTIA_CALL(NT) \
    if (nt == NT) return lres_mfma_inertia_##NT(wins, ids, nids, SW, HW, sigma, nshift, counts, total, skip, lin, st);
    LRES_MFMA_INERTIA_PARTS(LRES_INERTIA_CALL)
#undef LRES_INERTIA_CALL
    return hipErrorInvalidValue;
}

int mfma_scratch_bytes(int fp32, int nt, int nwc) {
    if (!mfma_nwc_ok(fp32, nt, nwc)) return -1;
#define LRES_MFMA_CALL(r, NT)                                                                                  \
    if (nt == NT && (fp32 != 0) == (sizeof(r) == 4))                                                           \
        return (int)lres_mfma_part_##r##_##NT(nwc == 0 ? -100 : -nwc, nullptr, nullptr, 0, nullptr, nullptr, 0, \
                                              0, nullptr, nullptr, nullptr, 0, WinOut{}, nullptr);
    LRES_MFMA_PARTS(LRES_MFMA_CALL)
#undef LRES_MFMA_CALL
    return -1;
}

hipError_t launch_reduce(const WinDesc* wins, int nwin, int ngroup, const double* Dpart, const double* Fpart,
                         const float* slots, int nslot, const int8_t* by_pole, const double* Q, double* D, double* F,
                         double* NB, double* EB, const double* circ, hipStream_t st) {
    // 16 waves per window: enough loads in flight for the per-pole partials (16 float
    // pairs per entry)
    hipLaunchKernelGGL(k_reduce, dim3(nwin), dim3(1024), 0, st, wins, ngroup, Dpart, Fpart,
                       reinterpret_cast<const float2*>(slots), nslot, by_pole, Q, D, F, NB, EB, circ);
    return hipGetLastError();
}

hipError_t launch_mirror(const WinDesc* wins, int nwin, double* D, double* F, hipStream_t st) {
    hipLaunchKernelGGL(k_mirror, dim3(nwin), dim3(256), 0, st, wins, D, F, nullptr);
    return hipGetLastError();
}

hipError_t launch_density(const WinDesc* wins, int nwin, int kmax, const double* D, DensRows rows, double* dens,
                          const double* skip, hipStream_t st) {
    const size_t lds = (size_t)kmax * kmax * sizeof(double);
    // D_B in LDS within the 64 KB default of dynamic LDS
    const int stage = lds <= 56 * 1024 ? 1 : 0;
    if (rows.chk_off) {
        if (stage)
            hipLaunchKernelGGL(k_density_chunk<true>, dim3(nwin), dim3(kDensBlkThreads), lds, st, wins, D, rows.own_g,
                               rows.chk_off, rows.chk_bs, rows.chk_idx, rows.chk_val, dens, skip);
        else
            hipLaunchKernelGGL(k_density_chunk<false>, dim3(nwin), dim3(kDensBlkThreads), 0, st, wins, D, rows.own_g,
                               rows.chk_off, rows.chk_bs, rows.chk_idx, rows.chk_val, dens, skip);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_mirror, dim3(nwin), dim3(256), 0, st, wins, const_cast<double*>(D), nullptr, skip);
    if (stage)
        hipLaunchKernelGGL(k_density<true>, dim3(nwin), dim3(64 * kDensWaves), lds, st, wins, D, rows.own_g, rows.xptr,
                           rows.xidx, rows.xval, dens, skip);
    else
        hipLaunchKernelGGL(k_density<false>, dim3(nwin), dim3(64 * kDensWaves), 0, st, wins, D, rows.own_g, rows.xptr,
                           rows.xidx, rows.xval, dens, skip);
    return hipGetLastError();
}

// Voronoi owner of every grid point (lrft/lcao3d.py:141-151): the nearest cell
// centre under the minimum image d - L rint(d / L) (round half to even), ties to
// the lowest cell (strict '<' in ascending order).  Every operation is written
// as a non-contracting intrinsic in the host builder's order, so the distances,
// and with them the ties, are bitwise those of the host code.
__global__ void __launch_bounds__(256) k_voronoi_owner(const double* __restrict__ cells, int ncell,
                                                       double L0, double L1, double L2, int n1, int n2, int n3,
                                                       double dx, int32_t* __restrict__ owner) {
    extern __shared__ double cs[];
    for (int q = threadIdx.x; q < 3 * ncell; q += blockDim.x) cs[q] = cells[q];
    __syncthreads();
    const int64_t ng = (int64_t)n1 * n2 * n3;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int64_t i = g / ((int64_t)n2 * n3), j = (g / n3) % n2, k = g % n3;
    const double r0 = __dmul_rn((double)i, dx), r1 = __dmul_rn((double)j, dx), r2 = __dmul_rn((double)k, dx);
    double best = 1.0 / 0.0;
    int bi = 0;
    for (int B = 0; B < ncell; ++B) {
        double d0 = __dsub_rn(r0, cs[3 * B + 0]);
        double d1 = __dsub_rn(r1, cs[3 * B + 1]);
        double d2 = __dsub_rn(r2, cs[3 * B + 2]);
        d0 = __dsub_rn(d0, __dmul_rn(L0, rint(__ddiv_rn(d0, L0))));
        d1 = __dsub_rn(d1, __dmul_rn(L1, rint(__ddiv_rn(d1, L1))));
        d2 = __dsub_rn(d2, __dmul_rn(L2, rint(__ddiv_rn(d2, L2))));
        const double sq = __dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)), __dmul_rn(d2, d2));
        if (sq < best) {
            best = sq;
            bi = B;
        }
    }
    owner[g] = bi;
}

hipError_t launch_voronoi_owner(const double* cells, int ncell, const double* box, const int64_t* shape, double dx,
                                int32_t* owner, hipStream_t st) {
    const int64_t ng = shape[0] * shape[1] * shape[2];
    const size_t lds = (size_t)3 * ncell * sizeof(double);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_voronoi_owner, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_voronoi_owner, dim3((unsigned)((ng + 255) / 256)), dim3(256), lds, st, cells, ncell, box[0],
                       box[1], box[2], (int)shape[0], (int)shape[1], (int)shape[2], dx, owner);
    return hipGetLastError();
}

// per window the smallest Rayleigh quotient of a basis function, min_a H_aa / S_aa
// (an upper bound of the window's lowest generalized eigenvalue)
__global__ void __launch_bounds__(64) k_diag_ratio(const WinDesc* __restrict__ wins, const double* __restrict__ SW,
                                                   const double* __restrict__ HW, double* __restrict__ out) {
    const int w = blockIdx.x;
    const WinDesc d = wins[w];
    double v = 1.0 / 0.0;
    for (int a = threadIdx.x; a < d.n; a += 64) {
        const int64_t q = d.mat_off + (int64_t)a * d.n + a;
        v = fmin(v, HW[q] / SW[q]);
    }
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_down(v, o, 64));
    if (threadIdx.x == 0) out[w] = v;
}

hipError_t launch_diag_ratio(const WinDesc* wins, int nwin, const double* SW, const double* HW, double* out,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_diag_ratio, dim3(nwin), dim3(64), 0, st, wins, SW, HW, out);
    return hipGetLastError();
}

hipError_t launch_set_circle(double* circ, double c, double R, double skip, double skip_density, hipStream_t st) {
    hipLaunchKernelGGL(k_set_circle, dim3(1), dim3(1), 0, st, circ, c, R, skip, skip_density);
    return hipGetLastError();
}

hipError_t launch_lmin_probe(const double* rq, int nwin, LminState* s, double tol, double* sig,
                             unsigned long long* itot, hipStream_t st) {
    hipLaunchKernelGGL(k_lmin_probe, dim3(1), dim3(256), 0, st, rq, nwin, s, tol, sig, itot);
    return hipGetLastError();
}

hipError_t launch_lmin_bracket(LminState* s, double* sig, unsigned long long* itot, const int32_t* ids, int nwin,
                               int32_t* cand, hipStream_t st) {
    hipLaunchKernelGGL(k_lmin_bracket, dim3(1), dim3(1024), 0, st, s, sig, itot, ids, nwin, cand);
    return hipGetLastError();
}

hipError_t launch_lmin_round(LminState* s, double* sig, unsigned long long* itot, const int32_t* counts, int kcap,
                             int32_t* cand, int32_t* tmp, hipStream_t st) {
    hipLaunchKernelGGL(k_lmin_round, dim3(1), dim3(1024), 0, st, s, sig, itot, counts, kcap, cand, tmp);
    return hipGetLastError();
}

hipError_t launch_lmin_snap(LminState* s, double width, double* sig, unsigned long long* itot, int* skip,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_lmin_snap, dim3(1), dim3(1), 0, st, s, width, sig, itot, skip);
    return hipGetLastError();
}

hipError_t launch_lmin_snap_end(LminState* s, double width, const double* sig, const unsigned long long* itot,
                                const int* skip, hipStream_t st) {
    hipLaunchKernelGGL(k_lmin_snap_end, dim3(1), dim3(1), 0, st, s, width, sig, itot, skip);
    return hipGetLastError();
}

hipError_t launch_edge_delta(const double* v, double* vcur, const double* vref, int64_t ng, double* part,
                             double* edge, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(kEdgeBlocks, (ng + 255) / 256);
    hipLaunchKernelGGL(k_edge_delta, dim3(nb), dim3(256), 0, st, v, vcur, vref, ng, part);
    hipLaunchKernelGGL(k_edge_final, dim3(1), dim3(256), 0, st, part, nb, edge);
    return hipGetLastError();
}

hipError_t launch_edge_ref(double* edge, double lo, double hi, hipStream_t st) {
    hipLaunchKernelGGL(k_edge_ref, dim3(1), dim3(1), 0, st, edge, lo, hi);
    return hipGetLastError();
}

hipError_t launch_step_edge(const double* prm, double* sig, unsigned long long* itot, double* edge, double safety,
                            int* skip, hipStream_t st) {
    hipLaunchKernelGGL(k_step_edge, dim3(1), dim3(1), 0, st, prm, sig, itot, edge, safety, skip);
    return hipGetLastError();
}

hipError_t launch_mu_begin(StepState* st, const unsigned long long* itot, const double* prm, double* circ,
                           hipStream_t stream) {
    hipLaunchKernelGGL(k_mu_begin, dim3(1), dim3(1), 0, stream, st, itot, prm, circ);
    return hipGetLastError();
}

hipError_t launch_mu_step(StepState* st, int stage, double* circ, void* pin, hipStream_t stream) {
    hipLaunchKernelGGL(k_mu_step, dim3(1), dim3(1), 0, stream, st, stage, circ,
                       reinterpret_cast<unsigned long long*>(pin));
    return hipGetLastError();
}

hipError_t launch_sum_step(const double* a, const double* b, int n, double* out, StepState* st, int stage,
                           double* circ, void* pin, hipStream_t stream) {
    hipLaunchKernelGGL(k_sum_step, dim3(1), dim3(256), 0, stream, a, b, n, out, st, stage, circ,
                       reinterpret_cast<unsigned long long*>(pin));
    return hipGetLastError();
}

hipError_t launch_sum2(const double* a, const double* b, int n, double* out, const double* skip, hipStream_t st) {
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(256), 0, st, a, b, n, out, skip);
    return hipGetLastError();
}

hipError_t launch_inertia(int T, int G, const WinDesc* wins, const int32_t* ids, int nids, const double* SW,
                          const double* HW, const double* sigma, int nshift, int32_t* counts,
                          unsigned long long* total, const int* skip, hipStream_t st) {
    if (G < 1 || G > kMaxG) return hipErrorInvalidValue;
    dim3 grid(nids, nshift), blk(G * G);
    switch (T) {
        case 2: hipLaunchKernelGGL((k_inertia<2>), grid, blk, 0, st, wins, ids, SW, HW, G, sigma, nshift, counts, total, skip); break;
        case 3: hipLaunchKernelGGL((k_inertia<3>), grid, blk, 0, st, wins, ids, SW, HW, G, sigma, nshift, counts, total, skip); break;
        case 4: hipLaunchKernelGGL((k_inertia<4>), grid, blk, 0, st, wins, ids, SW, HW, G, sigma, nshift, counts, total, skip); break;
        case 5: hipLaunchKernelGGL((k_inertia<5>), grid, blk, 0, st, wins, ids, SW, HW, G, sigma, nshift, counts, total, skip); break;
        case 6: hipLaunchKernelGGL((k_inertia<6>), grid, blk, 0, st, wins, ids, SW, HW, G, sigma, nshift, counts, total, skip); break;
        case 7: hipLaunchKernelGGL((k_inertia<7>), grid, blk, 0, st, wins, ids, SW, HW, G, sigma, nshift, counts, total, skip); break;
        case 8: hipLaunchKernelGGL((k_inertia<8>), grid, blk, 0, st, wins, ids, SW, HW, G, sigma, nshift, counts, total, skip); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// H(v) on the device (replaces LCAOBasis.H, lrft/lcao3d.py:98-100):
//   V_AB = dV sum_{r in supp A and supp B} v(r) x_A(r) x_B(r)^T   per atom pair A <= B,
// then H_W = T_W + V gathered into the window blocks.  The pair lists hold the
// common grid points of two atoms' supports in ascending order, so every V
// entry is summed in a fixed order (bitwise reproducible).
// ---------------------------------------------------------------------------
template <int NF>
__global__ void __launch_bounds__(64) k_pair_v(const int32_t* __restrict__ pairA, const int32_t* __restrict__ pairB,
                                               const int64_t* __restrict__ pent_off,
                                               const int32_t* __restrict__ ent_r, const int32_t* __restrict__ ent_a,
                                               const int32_t* __restrict__ ent_b,
                                               const int64_t* __restrict__ xat_off, const int32_t* __restrict__ nf_of,
                                               const double* __restrict__ xat, const int64_t* __restrict__ vp_off,
                                               const double* __restrict__ v, double dV, double* __restrict__ vpair) {
    const int p = blockIdx.x;
    const int A = pairA[p], B = pairB[p];
    const int nfa = nf_of[A], nfb = nf_of[B];
    const double* XA = xat + xat_off[A];
    const double* XB = xat + xat_off[B];
    double acc[NF][NF];
#pragma unroll
    for (int i = 0; i < NF; ++i)
#pragma unroll
        for (int j = 0; j < NF; ++j) acc[i][j] = 0.0;
    // Every load is issued unconditionally (function index clamped to the atom's
    // last, the value masked afterwards: a guarded load compiles to a branch and a
    // wait of its own), and the next entry's indices are loaded one step ahead.
    const int64_t e_end = pent_off[p + 1];
    int64_t e = pent_off[p] + threadIdx.x;
    const int64_t e0 = min(e, e_end - 1);
    int r_n = e_end > 0 ? ent_r[e0] : 0, a_n = e_end > 0 ? ent_a[e0] : 0, b_n = e_end > 0 ? ent_b[e0] : 0;
    for (; e < e_end; e += 64) {
        const int r = r_n, ia = a_n, ib = b_n;
        const int64_t en = min(e + 64, e_end - 1);
        r_n = ent_r[en];
        a_n = ent_a[en];
        b_n = ent_b[en];
        const double vr = v[r];
        const double* xa = XA + (int64_t)ia * nfa;
        const double* xb = XB + (int64_t)ib * nfb;
        double bv[NF], xv[NF];
#pragma unroll
        for (int j = 0; j < NF; ++j) bv[j] = xb[min(j, nfb - 1)];   // a paired atom has nf >= 1
#pragma unroll
        for (int i = 0; i < NF; ++i) xv[i] = xa[min(i, nfa - 1)];
#pragma unroll
        for (int j = 0; j < NF; ++j) bv[j] = j < nfb ? bv[j] : 0.0;
#pragma unroll
        for (int i = 0; i < NF; ++i) {
            const double av = i < nfa ? xv[i] * vr : 0.0;
#pragma unroll
            for (int j = 0; j < NF; ++j) acc[i][j] = fma(av, bv[j], acc[i][j]);
        }
    }
    double* out = vpair + vp_off[p];
#pragma unroll
    for (int i = 0; i < NF; ++i)
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            double t = acc[i][j];
            for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
            if (threadIdx.x == 0 && i < nfa && j < nfb) out[i * nfb + j] = t * dV;
        }
}

// H_W = T_W + V gathered into the window blocks (T-last order), over all windows'
// entries at once: hmap[q] is the place in vpair of entry q's V value, or -1 where
// the two functions' atoms share no grid point (built on the host with the pair
// tables: pairs are stored once with A <= B, the other orientation reads the
// transposed block).
__global__ void __launch_bounds__(256) k_window_hv(int64_t total, const int32_t* __restrict__ hmap,
                                                   const double* __restrict__ vpair, const double* __restrict__ TW,
                                                   double* __restrict__ HW) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int32_t m = hmap[q];
        const double t = TW[q];
        HW[q] = m >= 0 ? t + vpair[m] : t;
    }
}

hipError_t launch_pair_v(int nf_max, int npair, const int32_t* pairA, const int32_t* pairB, const int64_t* pent_off,
                         const int32_t* ent_r, const int32_t* ent_a, const int32_t* ent_b, const int64_t* xat_off,
                         const int32_t* nf_of, const double* xat, const int64_t* vp_off, const double* v, double dV,
                         double* vpair, hipStream_t st) {
    if (npair == 0) return hipSuccess;
    if (nf_max <= 6)
        hipLaunchKernelGGL(k_pair_v<6>, dim3(npair), dim3(64), 0, st, pairA, pairB, pent_off, ent_r, ent_a, ent_b,
                           xat_off, nf_of, xat, vp_off, v, dV, vpair);
    else if (nf_max <= 8)
        hipLaunchKernelGGL(k_pair_v<8>, dim3(npair), dim3(64), 0, st, pairA, pairB, pent_off, ent_r, ent_a, ent_b,
                           xat_off, nf_of, xat, vp_off, v, dV, vpair);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_window_hv(int64_t total, const int32_t* hmap, const double* vpair, const double* TW, double* HW,
                            hipStream_t st) {
    if (total == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_window_hv, dim3((unsigned)blocks), dim3(256), 0, st, total, hmap, vpair, TW, HW);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Windows larger than the register-resident sweep: blocked Gauss-Jordan with the
// matrix in a global workspace (one slot per block, reused across its poles).
// Block step with pivot block P = [p0, p0 + b): column panel C = A[:, P] and row
// panel R = A[P, :] go to LDS; Pi = A_PP^-1 is formed by one wave (scalar
// shifted-pivot sweep); then with C~ = C - [I on P] and R~ = Pi (R + [I on P])
// the uniform rank-b update A -= C~ R~ is the blocked Gauss-Jordan step (row
// block -> Pi R, column block -> -C Pi, pivot block -> Pi, rest A - C Pi R).
// ---------------------------------------------------------------------------
template <typename E> struct Fld;
template <> struct Fld<double> {
    __device__ static double zero() { return 0.0; }
    __device__ static double one() { return 1.0; }
    __device__ static double add_one(double a, double s) { return a + s; }
    __device__ static double mul(double a, double b) { return a * b; }
    __device__ static double fms(double acc, double a, double b) { return fma(-a, b, acc); }
    __device__ static double fma_(double acc, double a, double b) { return fma(a, b, acc); }
    __device__ static double recip(double a) { return 1.0 / a; }
    __device__ static bool positive(double a) { return a > 0.0; }
};
template <typename R> struct Fld<cx<R>> {
    __device__ static cx<R> zero() { return {R(0), R(0)}; }
    __device__ static cx<R> one() { return {R(1), R(0)}; }
    __device__ static cx<R> add_one(cx<R> a, R s) { return {a.re + s, a.im}; }
    __device__ static cx<R> mul(cx<R> a, cx<R> b) { return cmul(a, b); }
    __device__ static cx<R> fms(cx<R> acc, cx<R> a, cx<R> b) { return cupd(acc, -a.re, -a.im, a.im, b); }
    __device__ static cx<R> fma_(cx<R> acc, cx<R> a, cx<R> b) {
        return {fma(a.re, b.re, fma(-a.im, b.im, acc.re)), fma(a.re, b.im, fma(a.im, b.re, acc.im))};
    }
    __device__ static cx<R> recip(cx<R> a) { return crecip(a); }
    __device__ static bool positive(cx<R> a) { return a.re > R(0); }
};

__device__ inline double shfl_e(double v, int src) { return __shfl(v, src, 64); }
template <typename R> __device__ inline cx<R> shfl_e(cx<R> v, int src) {
    return {__shfl(v.re, src, 64), __shfl(v.im, src, 64)};
}

// Pi = A_PP^-1 of the b x b pivot block (b <= 16; rows p0.. of the row panel Rl,
// columns p0..) into Pl (b x b, row-major) by the scalar shifted-pivot sweep
// a_ij -= (a_ip - d_ip)(a_pj + d_pj) / a_pp, whose in-place result is the inverse.
// The block (padded to 16 x 16 with the identity, on which the sweep is a no-op) sits
// in wave 0's registers, lane l holding column l % 16 of rows 4 (l / 16) + r, r < 4;
// each pivot's row, column and entry come from lane shuffles, so the pivots run
// without LDS round trips.  Returns (in wave 0) the number of positive scalar pivots,
// which for the unpivoted block are the LDL^T pivots counted by the inertia.
template <typename E>
__device__ int big_pivot_inverse(const E* Rl, E* Pl, int n, int p0, int b) {
    using Fo = Fld<E>;
    if (threadIdx.x >= 64) return 0;
    const int lane = threadIdx.x;
    const int j = lane & 15, g = lane >> 4;
    E a[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = 4 * g + r;
        a[r] = (i < b && j < b) ? Rl[i * n + p0 + j] : Fo::add_one(Fo::zero(), i == j ? 1 : 0);
    }
    int npos = 0;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const E rowp = shfl_e(a[p & 3], (p >> 2) * 16 + j);   // A[p][j]
        const E piv = shfl_e(a[p & 3], (p >> 2) * 16 + p);    // A[p][p]
        E colp[4];                                             // A[4 g + r][p]
#pragma unroll
        for (int r = 0; r < 4; ++r) colp[r] = shfl_e(a[r], g * 16 + p);
        if (p < b && Fo::positive(piv)) ++npos;
        const E rv = Fo::mul((j == p) ? Fo::add_one(rowp, 1) : rowp, Fo::recip(piv));
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = Fo::fms(a[r], (4 * g + r == p) ? Fo::add_one(colp[r], -1) : colp[r], rv);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = 4 * g + r;
        if (i < b && j < b) Pl[i * b + j] = a[r];
    }
    return npos;
}

// 16 x 16 x 4 matrix-core products per real type.  Layouts (gfx950, checked with one-hot
// operands): A operand lane l -> A[l % 16][l / 16], B operand lane l -> B[l / 16][l % 16];
// accumulator lane l, register r -> column l % 16 and row l / 16 + 4 r (f64) or
// 4 (l / 16) + r (f32).
template <typename R> struct Mfma16;
template <> struct Mfma16<double> {
    typedef double V4 __attribute__((ext_vector_type(4)));
    __device__ static V4 op(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lg, int r) { return lg + 4 * r; }
};
template <> struct Mfma16<float> {
    typedef float V4 __attribute__((ext_vector_type(4)));
    __device__ static V4 op(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lg, int r) { return 4 * lg + r; }
};

// R~ = Pi (R + [I on P]) in place on the row panel's columns >= lo: one 16-column tile per
// wave on the matrix cores (a 16 x 16 by 16 x 16 complex product, four real MFMAs per
// k-step of four), operands padded with zeros beyond b; each wave reads all of its
// columns before writing them back.
template <typename R>
__device__ void big_rows_tilde(cx<R>* Rl, const cx<R>* Pl, int n, int lo, int p0, int b) {
    using Mf = Mfma16<R>;
    typedef typename Mf::V4 V4;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    for (int T = (lo >> 4) + wv; T < (n + 15) >> 4; T += nw) {
        const int col = 16 * T + li;
        const bool colok = col >= lo && col < n;
        const cx<R> zero = {R(0), R(0)};
        cx<R> bv[4], av[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int q = 4 * kk + lg;
            cx<R> v = (q < b && colok) ? Rl[q * n + col] : zero;
            if (q < b && col == p0 + q) v.re += R(1);
            bv[kk] = v;
            av[kk] = (li < b && q < b) ? Pl[li * b + q] : zero;
        }
        V4 re = {R(0), R(0), R(0), R(0)}, im = {R(0), R(0), R(0), R(0)};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            re = Mf::op(av[kk].re, bv[kk].re, re);
            re = Mf::op(-av[kk].im, bv[kk].im, re);
            im = Mf::op(av[kk].re, bv[kk].im, im);
            im = Mf::op(av[kk].im, bv[kk].re, im);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = Mf::row(lg, r);
            if (q < b && colok) Rl[q * n + col] = cx<R>{re[r], im[r]};
        }
    }
}

// real fp64: one v_mfma_f64_16x16x4f64 per k-step of four, tiles as above
__device__ void big_rows_tilde(double* Rl, const double* Pl, int n, int lo, int p0, int b) {
    typedef double V4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    for (int T = (lo >> 4) + wv; T < (n + 15) >> 4; T += nw) {
        const int col = 16 * T + li;
        const bool colok = col >= lo && col < n;
        V4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int q = 4 * kk + lg;
            double v = (q < b && colok) ? Rl[q * n + col] : 0.0;
            if (q < b && col == p0 + q) v += 1.0;
            const double pa = (li < b && q < b) ? Pl[li * b + q] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pa, v, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = lg + 4 * r;
            if (q < b && colok) Rl[q * n + col] = acc[r];
        }
    }
}

// A -= C~ R~ on lo <= i <= j (C~ stored transposed, Cl[q n + i]; R~ in Rl[q n + j]) on
// the matrix cores: one 16 x 16 tile (I <= J) of the upper triangle per wave at a time,
// four real MFMAs per complex product and k-step of four panel columns, the next tile's
// entries loaded while this tile's products run.  Operand rows / columns beyond n read
// as zero; rows or columns below lo (not kept in the panels) only reach output entries
// below lo, which are not stored.
template <typename R>
__device__ void big_update_sym(cx<R>* __restrict__ A, int n, int lo, int b, const cx<R>* Cl, const cx<R>* Rl) {
    using Mf = Mfma16<R>;
    typedef typename Mf::V4 V4;
    const cx<R> zero = {R(0), R(0)};
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int t0 = lo >> 4, ntile = ((n + 15) >> 4) - t0;
    const int tiles = ntile * (ntile + 1) / 2;
    const int nk = (b + 3) >> 2;
    // tile t -> (I, J), I <= J, row-major over the upper triangle of the tile grid
    auto tile_of = [&](int t, int& r0, int& c0) {
        int I = 0, rem = t;
        while (rem >= ntile - I) {
            rem -= ntile - I;
            ++I;
        }
        r0 = 16 * (t0 + I);
        c0 = 16 * (t0 + I + rem);
    };
    auto load = [&](int r0, int c0, V4& re, V4& im) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + Mf::row(lg, r), col = c0 + li;
            const bool ok = row >= lo && row <= col && col < n;
            const cx<R> a = ok ? A[(int64_t)row * n + col] : zero;
            re[r] = a.re;
            im[r] = a.im;
        }
    };
    int t = wv;
    if (t >= tiles) return;
    int r0, c0;
    tile_of(t, r0, c0);
    V4 are, aim;
    load(r0, c0, are, aim);
    for (; t < tiles; t += nw) {
        // the next tile's entries are loaded while this tile's products run
        int r1 = 0, c1 = 0;
        V4 bre = {R(0), R(0), R(0), R(0)}, bim = {R(0), R(0), R(0), R(0)};
        const bool more = t + nw < tiles;
        if (more) {
            tile_of(t + nw, r1, c1);
            load(r1, c1, bre, bim);
        }
        for (int kk = 0; kk < nk; ++kk) {
            const int q = 4 * kk + lg;
            const int ra = r0 + li, cb = c0 + li;
            const cx<R> ca = (q < b && ra < n) ? Cl[q * n + ra] : zero;
            const cx<R> rb = (q < b && cb < n) ? Rl[q * n + cb] : zero;
            // A -= C R:  re += (-Cr) Rr + Ci Ri,  im += (-Cr) Ri + (-Ci) Rr
            are = Mf::op(-ca.re, rb.re, are);
            are = Mf::op(ca.im, rb.im, are);
            aim = Mf::op(-ca.re, rb.im, aim);
            aim = Mf::op(-ca.im, rb.re, aim);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + Mf::row(lg, r), col = c0 + li;
            if (row >= lo && row <= col && col < n) A[(int64_t)row * n + col] = cx<R>{are[r], aim[r]};
        }
        if (!more) break;
        r0 = r1;
        c0 = c1;
        are = bre;
        aim = bim;
    }
}

// real fp64: the same tiles with one v_mfma_f64_16x16x4f64 per k-step (the inertia
// counts' LDL^T of the real symmetric sigma S_W - H_W)
__device__ void big_update_sym(double* __restrict__ A, int n, int lo, int b, const double* Cl, const double* Rl) {
    typedef double V4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int t0 = lo >> 4, ntile = ((n + 15) >> 4) - t0;
    const int tiles = ntile * (ntile + 1) / 2;
    const int nk = (b + 3) >> 2;
    auto tile_of = [&](int t, int& r0, int& c0) {
        int I = 0, rem = t;
        while (rem >= ntile - I) {
            rem -= ntile - I;
            ++I;
        }
        r0 = 16 * (t0 + I);
        c0 = 16 * (t0 + I + rem);
    };
    auto load = [&](int r0, int c0, V4& acc) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + lg + 4 * r, col = c0 + li;
            acc[r] = (row >= lo && row <= col && col < n) ? A[(int64_t)row * n + col] : 0.0;
        }
    };
    int t = wv;
    if (t >= tiles) return;
    int r0, c0;
    tile_of(t, r0, c0);
    V4 acc;
    load(r0, c0, acc);
    for (; t < tiles; t += nw) {
        int r1 = 0, c1 = 0;
        V4 nxt = {0.0, 0.0, 0.0, 0.0};
        const bool more = t + nw < tiles;
        if (more) {
            tile_of(t + nw, r1, c1);
            load(r1, c1, nxt);
        }
        for (int kk = 0; kk < nk; ++kk) {
            const int q = 4 * kk + lg;
            const int ra = r0 + li, cb = c0 + li;
            const double ca = (q < b && ra < n) ? Cl[q * n + ra] : 0.0;
            const double rb = (q < b && cb < n) ? Rl[q * n + cb] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-ca, rb, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + lg + 4 * r, col = c0 + li;
            if (row >= lo && row <= col && col < n) A[(int64_t)row * n + col] = acc[r];
        }
        if (!more) break;
        r0 = r1;
        c0 = c1;
        acc = nxt;
    }
}

// One blocked Gauss-Jordan step on the n x n matrix A (global, row-major).
// LDS: Cl (n x b), Rl (b x n), Pl (b x b), b <= 16.  Returns (in wave 0) the number of
// positive scalar pivots met inside the pivot block.
template <typename E>
__device__ int big_block_step(E* __restrict__ A, int n, int p0, int b, E* Cl, E* Rl, E* Pl, int lo = 0) {
    using Fo = Fld<E>;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < n * b; idx += nt) {
        const int i = idx / b, q = idx - (idx / b) * b;
        Cl[idx] = A[(int64_t)i * n + p0 + q];
    }
    for (int idx = tid; idx < b * n; idx += nt) Rl[idx] = A[(int64_t)p0 * n + idx];
    __syncthreads();
    const int npos = big_pivot_inverse(Rl, Pl, n, p0, b);   // Pi = A_PP^-1 into Pl
    __syncthreads();
    big_rows_tilde(Rl, Pl, n, lo, p0, b);   // R~ = Pi (R + [I on P]) in place; C~ = C - [I on P]
    for (int q = tid; q < b; q += nt) Cl[(p0 + q) * b + q] = Fo::add_one(Cl[(p0 + q) * b + q], -1);
    __syncthreads();
    // A -= C~ R~ on rows and columns >= lo: each thread four rows of one column per round
    // (row-major, coalesced in j).  lo = 0: the Gauss-Jordan step of the whole matrix
    const int nl = n - lo;
    const int rows4 = (nl + 3) / 4;
    for (int idx = tid; idx < rows4 * nl; idx += nt) {
        const int i0 = lo + (idx / nl) * 4, j = lo + idx - (idx / nl) * nl;
        E a[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) a[s] = (i0 + s < n) ? A[(int64_t)(i0 + s) * n + j] : Fo::zero();
        for (int q = 0; q < b; ++q) {
            const E r = Rl[q * n + j];
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (i0 + s < n) a[s] = Fo::fms(a[s], Cl[(i0 + s) * b + q], r);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (i0 + s < n) A[(int64_t)(i0 + s) * n + j] = a[s];
    }
    __syncthreads();
    return npos;
}

// big_block_step on the upper triangle only, for complex-symmetric A (the pole
// solves).  After the pivots of a set D the in-place state is symmetric on D x D
// and on the rest U x U and antisymmetric between them (W[u][d] = -W[d][u]), so
// entries (i, j) with i <= j carry everything: the pivot rows are read from the
// upper triangle (A[p][j] = -A[j][p] for pivoted j < p0, the mirror inside the
// pivot block), the pivot columns are the pivot rows with the sign flipped on the
// pivoted rows, and the update touches i <= j only (half the workspace traffic).
// On return the upper triangle holds that of the Gauss-Jordan state.
//
// Only rows and columns >= lo are read and written.  The caller needs the block
// [m, n) x [m, n) of the inverse (the T-last window order puts T_B there): while the
// pivots lie below m, lo = p0 makes the step a right-looking elimination of the
// trailing block (after the pivots of D = [0, m) it holds the Schur complement
// A_TT - A_TD A_DD^-1 A_DT, and rows / columns of D are never read again); from m on,
// lo = m runs Gauss-Jordan on that k x k block, whose in-place result is its inverse,
// (A^-1)[T,T].  With lo = 0 this is the whole-matrix sweep.  The same steps with lo = p0
// throughout are the unpivoted LDL^T of a real symmetric A; the return value (in wave 0)
// counts the positive pivots of the block.
template <typename E>
__device__ int big_block_step_sym(E* __restrict__ A, int n, int p0, int b, E* Cl, E* Rl, E* Pl, int lo) {
    using Fo = Fld<E>;
    const int tid = threadIdx.x, nt = blockDim.x;
    // pivot rows: right of the diagonal from the rows themselves (coalesced along j),
    // left of it from the columns, read as rows j of the b pivot columns (contiguous)
    for (int idx = tid; idx < b * n; idx += nt) {
        const int q = idx / n, j = idx - (idx / n) * n;
        if (j >= p0 + q) Rl[idx] = A[(int64_t)(p0 + q) * n + j];
    }
    for (int idx = tid; idx < (p0 + b) * b; idx += nt) {
        const int j = idx / b, q = idx - (idx / b) * b;
        if (j >= lo && j < p0 + q) {
            const E v = A[(int64_t)j * n + p0 + q];
            Rl[q * n + j] = j < p0 ? Fo::mul(v, Fo::add_one(Fo::zero(), -1)) : v;
        }
    }
    __syncthreads();
    // pivot columns (stored transposed, Cl[q n + i]: consecutive rows at consecutive
    // addresses for the matrix-core operands): the pivot rows with the sign flipped on
    // the pivoted rows
    for (int idx = tid; idx < n * b; idx += nt) {
        const int q = idx / n, i = idx - (idx / n) * n;
        if (i < lo) continue;
        const E v = Rl[q * n + i];
        Cl[idx] = i < p0 ? Fo::mul(v, Fo::add_one(Fo::zero(), -1)) : v;
    }
    __syncthreads();
    const int npos = big_pivot_inverse(Rl, Pl, n, p0, b);   // Pi = A_PP^-1 into Pl
    __syncthreads();
    big_rows_tilde(Rl, Pl, n, lo, p0, b);  // R~ = Pi (R + [I on P]) in place on columns >= lo
    for (int q = tid; q < b; q += nt) Cl[q * n + p0 + q] = Fo::add_one(Cl[q * n + p0 + q], -1);
    __syncthreads();
    big_update_sym(A, n, lo, b, Cl, Rl);
    __syncthreads();
    return npos;
}

constexpr int kBigThreads = 1024;

template <typename R>
__global__ void __launch_bounds__(kBigThreads) k_pole_big(const WinDesc* __restrict__ wins, const int32_t* __restrict__ ids,
                                                  int nids, const double* __restrict__ SW, const double* __restrict__ HW,
                                                  int M, int ngroup, const double* __restrict__ circ, int bmax,
                                                  cx<double>* __restrict__ work, const int64_t* __restrict__ ws_off,
                                                  double* __restrict__ Dpart, double* __restrict__ Fpart) {
    extern __shared__ double lds_raw[];
    using E = cx<R>;
    if (circ[2] != 0.0) return;   // a trial pass the device-side mu search did not need
    const double c = circ[0], Rr = circ[1];
    // XCD-aware order (workgroups go to the 8 XCDs round-robin by id): every pole group
    // of a window on one XCD, dispatched together, so S_W and H_W come from its L2
    const int L = blockIdx.x;
    const int slot = (L / (8 * ngroup)) * 8 + (L & 7);
    const int g = (L >> 3) % ngroup;
    if (slot >= nids) return;
    const int w = ids[slot];
    const WinDesc d = wins[w];
    const int n = d.n, m = d.m, k = d.k;
    E* A = reinterpret_cast<E*>(work + ws_off[w] + (int64_t)g * n * n);
    E* Cl = reinterpret_cast<E*>(lds_raw);
    E* Rl = Cl + (size_t)n * bmax;
    E* Pl = Rl + (size_t)n * bmax;
    const double* S = SW + d.mat_off;
    const double* H = HW + d.mat_off;
    double* dp = Dpart + d.kk_off * ngroup + (int64_t)g * k * k;
    double* fp = Fpart + d.kk_off * ngroup + (int64_t)g * k * k;
    bool first = true;
    for (int jp = g; jp < M / 2; jp += ngroup) {
        cx<double> z, wt, wz;
        pole(jp, M, c, Rr, z, wt, wz);
        for (int64_t q = threadIdx.x; q < (int64_t)n * n; q += blockDim.x) {
            if (q / n > q % n) continue;   // upper triangle only
            A[q] = {(R)(z.re * S[q] - H[q]), (R)(z.im * S[q])};
        }
        __syncthreads();
        // pivots below m: elimination of the trailing block; from m on: Gauss-Jordan on T x T
        for (int p0 = 0; p0 < n; p0 += bmax) {
            const int bb = min(bmax, n - p0);
            if (p0 < m && p0 + bb > m) {   // a block straddling m: split it there
                big_block_step_sym<E>(A, n, p0, m - p0, Cl, Rl, Pl, p0);
                big_block_step_sym<E>(A, n, m, p0 + bb - m, Cl, Rl, Pl, m);
            } else {
                big_block_step_sym<E>(A, n, p0, bb, Cl, Rl, Pl, p0 < m ? p0 : m);
            }
        }
        for (int q = threadIdx.x; q < k * k; q += blockDim.x) {
            const int r = q / k, col = q - (q / k) * k;
            if (r > col) continue;   // upper triangle (the partial sums are symmetric)
            const E a = A[(int64_t)(r + m) * n + col + m];
            const double gr = (double)a.re, gi = (double)a.im;
            const double dv = wt.re * gr - wt.im * gi, fv = wz.re * gr - wz.im * gi;
            if (first) {
                dp[q] = dv;
                fp[q] = fv;
            } else {
                dp[q] += dv;
                fp[q] += fv;
            }
        }
        __syncthreads();
        first = false;
    }
}

// Inertia counts of sigma_s S_W - H_W by the unpivoted LDL^T (right-looking blocked
// elimination on the upper triangle, real fp64 on the matrix cores).  Block (x, y)
// factors shifts y, y + gridDim.y, ... of window ids[x] in its own slot of the
// workspace (stride entries apart), so every (window, shift) of the launch can run at
// once when the workspace holds that many slots.
__global__ void __launch_bounds__(kBigThreads) k_inertia_big(const WinDesc* __restrict__ wins, const int32_t* __restrict__ ids,
                                                     const double* __restrict__ SW, const double* __restrict__ HW,
                                                     const double* __restrict__ sigma, int nshift, int bmax,
                                                     double* __restrict__ work, int64_t stride,
                                                     int32_t* __restrict__ counts,
                                                     unsigned long long* __restrict__ total,
                                                     const int* __restrict__ skip) {
    extern __shared__ double lds_raw[];
    if (skip != nullptr && *skip != 0) return;   // decided by the keep-edge bound
    const int w = ids[blockIdx.x];
    const WinDesc d = wins[w];
    const int n = d.n;
    double* A = work + ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) * stride;
    double* Cl = lds_raw;
    double* Rl = Cl + (size_t)n * bmax;
    double* Pl = Rl + (size_t)n * bmax;
    const double* S = SW + d.mat_off;
    const double* H = HW + d.mat_off;
    for (int s_id = blockIdx.y; s_id < nshift; s_id += gridDim.y) {
        const double sg = sigma[s_id];
        for (int64_t q = threadIdx.x; q < (int64_t)n * n; q += blockDim.x) {
            if (q / n > q % n) continue;   // upper triangle only
            A[q] = sg * S[q] - H[q];
        }
        __syncthreads();
        int npos = 0;
        for (int p0 = 0; p0 < n; p0 += bmax)
            npos += big_block_step_sym<double>(A, n, p0, min(bmax, n - p0), Cl, Rl, Pl, p0);
        if (threadIdx.x == 0) {
            if (counts) counts[(int64_t)w * nshift + s_id] = npos;
            atomicAdd(&total[s_id], (unsigned long long)npos);
        }
        __syncthreads();
    }
}

// dynamic LDS of the big kernels: two panels of n x b and the b x b pivot block
size_t big_lds_bytes(int nmax, int b, int elem_bytes) {
    return ((size_t)2 * nmax * b + (size_t)b * b) * elem_bytes;
}

hipError_t launch_pole_big(int fp32, int bmax, int nmax, const WinDesc* wins, const int32_t* ids, int nids,
                           const double* SW, const double* HW, int M, int ngroup, const double* circ, void* work,
                           const int64_t* ws_off, double* Dpart, double* Fpart, hipStream_t st) {
    const size_t lds = big_lds_bytes(nmax, bmax, fp32 ? 8 : 16);
    dim3 grid((nids + 7) / 8 * 8 * ngroup), blk(kBigThreads);   // ids past nids exit at once
    hipError_t e;
    if (fp32) {
        e = hipFuncSetAttribute((const void*)k_pole_big<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_pole_big<float>, grid, blk, lds, st, wins, ids, nids, SW, HW, M, ngroup, circ, bmax,
                           (cx<double>*)work, ws_off, Dpart, Fpart);
    } else {
        e = hipFuncSetAttribute((const void*)k_pole_big<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_pole_big<double>, grid, blk, lds, st, wins, ids, nids, SW, HW, M, ngroup, circ, bmax,
                           (cx<double>*)work, ws_off, Dpart, Fpart);
    }
    return hipGetLastError();
}

hipError_t launch_inertia_big(int bmax, int nmax, const WinDesc* wins, const int32_t* ids, int nids, const double* SW,
                              const double* HW, const double* sigma, int nshift, double* work, int64_t work_len,
                              int32_t* counts, unsigned long long* total, const int* skip, hipStream_t st) {
    const size_t lds = big_lds_bytes(nmax, bmax, 8);
    const int64_t stride = (int64_t)nmax * nmax;
    const int64_t slots = work_len / stride;
    if (nids <= 0 || slots < nids) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute((const void*)k_inertia_big, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int ny = (int)std::min<int64_t>(nshift, slots / nids);
    dim3 grid(nids, ny), blk(kBigThreads);
    hipLaunchKernelGGL(k_inertia_big, grid, blk, lds, st, wins, ids, SW, HW, sigma, nshift, bmax, work, stride, counts,
                       total, skip);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Adjoint of the window step at a fixed circle (mu, e_lo).  With the weight
// p(z) = alpha z + beta per pole and G_j = (z_j S_W - H_W)^-1,
//   d/dH_W sum_B (alpha E_B + beta N_B) =  2 Re sum_j w_j p(z_j)     G_j[:,T] Q_B G_j[T,:]
//   d/dS_W sum_B (alpha E_B + beta N_B) = -2 Re sum_j w_j z_j p(z_j) G_j[:,T] Q_B G_j[T,:]
// (E_B = Tr(F_B Q_B), N_B = Tr(D_B Q_B); dG = G dH G - z G dS G; G and Q are
// symmetric).  Three kernels per chunk of windows: every G_j by the blocked
// in-place Gauss-Jordan sweep (fp64) into a workspace, Z_j = Q_B G_j[T,:], and
// the pole sum of G_j[:,T] Z_j accumulated tile by tile in registers, poles in
// order (bitwise reproducible).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBigThreads) k_adj_invert(const WinDesc* __restrict__ wins,
                                                            const int32_t* __restrict__ ids,
                                                            const int64_t* __restrict__ woff,
                                                            const double* __restrict__ SW,
                                                            const double* __restrict__ HW, int M, double c,
                                                            double Rr, int bmax, cx<double>* __restrict__ Gw) {
    extern __shared__ double lds_raw[];
    using E = cx<double>;
    const int w = ids[blockIdx.x];
    const int j = blockIdx.y;
    const WinDesc d = wins[w];
    const int n = d.n;
    E* A = Gw + woff[blockIdx.x] + (int64_t)j * n * n;
    E* Cl = reinterpret_cast<E*>(lds_raw);
    E* Rl = Cl + (size_t)n * bmax;
    E* Pl = Rl + (size_t)n * bmax;
    const double* S = SW + d.mat_off;
    const double* H = HW + d.mat_off;
    cx<double> z, wt, wz;
    pole(j, M, c, Rr, z, wt, wz);
    for (int64_t q = threadIdx.x; q < (int64_t)n * n; q += blockDim.x) A[q] = {z.re * S[q] - H[q], z.im * S[q]};
    __syncthreads();
    for (int p0 = 0; p0 < n; p0 += bmax) big_block_step<E>(A, n, p0, min(bmax, n - p0), Cl, Rl, Pl);
}

// Z_j = Q_B G_j[T,:] (k x n) into a second workspace of G's shape
__global__ void __launch_bounds__(256) k_adj_qg(const WinDesc* __restrict__ wins, const int32_t* __restrict__ ids,
                                                const int64_t* __restrict__ woff, const double* __restrict__ Q,
                                                const cx<double>* __restrict__ Gw, cx<double>* __restrict__ Zw) {
    const int w = ids[blockIdx.x];
    const int j = blockIdx.y;
    const WinDesc d = wins[w];
    const int n = d.n, m = d.m, k = d.k;
    const cx<double>* G = Gw + woff[blockIdx.x] + (int64_t)j * n * n;
    cx<double>* Z = Zw + woff[blockIdx.x] + (int64_t)j * n * n;
    const double* Qb = Q + d.kk_off;
    for (int64_t e = threadIdx.x; e < (int64_t)k * n; e += blockDim.x) {
        const int t = (int)(e / n), b = (int)(e - (int64_t)t * n);
        double re = 0.0, im = 0.0;
        for (int u = 0; u < k; ++u) {
            const double q = Qb[(int64_t)t * k + u];
            const cx<double> g = G[(int64_t)(m + u) * n + b];
            re = fma(q, g.re, re);
            im = fma(q, g.im, im);
        }
        Z[e] = {re, im};
    }
}

// gH, gS (n x n per window, the plan's T-last layout): one 16 x 16 output tile
// per block, the pole sum in registers
__global__ void __launch_bounds__(256) k_adj_accum(const WinDesc* __restrict__ wins, const int32_t* __restrict__ ids,
                                                   const int64_t* __restrict__ woff,
                                                   const cx<double>* __restrict__ Gw,
                                                   const cx<double>* __restrict__ Zw, int M, double c, double Rr,
                                                   double alpha, double beta, double* __restrict__ gH,
                                                   double* __restrict__ gS) {
    __shared__ cx<double> Gs[16][17];
    __shared__ cx<double> Zs[16][17];
    const int w = ids[blockIdx.x];
    const WinDesc d = wins[w];
    const int n = d.n, m = d.m, k = d.k;
    const int nt = (n + 15) / 16;
    const int TI = blockIdx.y / nt, TJ = blockIdx.y - (blockIdx.y / nt) * nt;
    if (TI >= nt) return;   // uniform per block
    const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
    const int a = 16 * TI + ty, b = 16 * TJ + tx;
    double sh = 0.0, ss = 0.0;
    for (int j = 0; j < M / 2; ++j) {
        const cx<double>* G = Gw + woff[blockIdx.x] + (int64_t)j * n * n;
        const cx<double>* Z = Zw + woff[blockIdx.x] + (int64_t)j * n * n;
        double yr = 0.0, yi = 0.0;
        for (int t0 = 0; t0 < k; t0 += 16) {
            const int tg = t0 + tx, tz = t0 + ty;
            Gs[ty][tx] = (a < n && tg < k) ? G[(int64_t)a * n + m + tg] : cx<double>{0.0, 0.0};
            Zs[ty][tx] = (tz < k && b < n) ? Z[(int64_t)tz * n + b] : cx<double>{0.0, 0.0};
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const cx<double> g = Gs[ty][q], zz = Zs[q][tx];
                yr = fma(g.re, zz.re, fma(-g.im, zz.im, yr));
                yi = fma(g.re, zz.im, fma(g.im, zz.re, yi));
            }
            __syncthreads();
        }
        cx<double> z, wt, wz;
        pole(j, M, c, Rr, z, wt, wz);
        // coefficients 2 w p(z) and -2 w z p(z), p(z) = alpha z + beta
        const cx<double> p = {alpha * z.re + beta, alpha * z.im};
        const cx<double> ch = cmul(wt, p);
        const cx<double> cs = cmul(ch, z);
        sh += 2.0 * (ch.re * yr - ch.im * yi);
        ss -= 2.0 * (cs.re * yr - cs.im * yi);
    }
    if (a < n && b < n) {
        gH[d.mat_off + (int64_t)a * n + b] = sh;
        gS[d.mat_off + (int64_t)a * n + b] = ss;
    }
}

hipError_t launch_adjoint(int bmax, int nmax, const WinDesc* wins, const int32_t* ids, const int64_t* woff, int nids,
                          const double* SW, const double* HW, const double* Q, int M, double c, double Rr,
                          double alpha, double beta, void* Gw, void* Zw, double* gH, double* gS, hipStream_t st) {
    const size_t lds = big_lds_bytes(nmax, bmax, 16);
    hipError_t e = hipFuncSetAttribute((const void*)k_adj_invert, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_invert, dim3(nids, M / 2), dim3(kBigThreads), lds, st, wins, ids, woff, SW, HW, M, c, Rr,
                       bmax, (cx<double>*)Gw);
    hipLaunchKernelGGL(k_adj_qg, dim3(nids, M / 2), dim3(256), 0, st, wins, ids, woff, Q, (const cx<double>*)Gw,
                       (cx<double>*)Zw);
    const int nt = (nmax + 15) / 16;
    hipLaunchKernelGGL(k_adj_accum, dim3(nids, nt * nt), dim3(256), 0, st, wins, ids, woff, (const cx<double>*)Gw,
                       (const cx<double>*)Zw, M, c, Rr, alpha, beta, gH, gS);
    return hipGetLastError();
}
