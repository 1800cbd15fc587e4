// Pole solves on the matrix cores (MFMA) for windows of up to 128 functions (both
// precisions) and, in fp32, up to 256.
//
// One workgroup of NWC waves handles one (window, pole group).  Per pole it
// holds A = z S_W - H_W, padded to 16 NT x 16 NT (zero off-diagonal, 1 on the
// padding diagonal), as 16 x 16 tiles in MFMA accumulator registers.  The
// Gauss-Jordan sweep runs in blocks of four pivots P = {4p, .., 4p + 3}.  With
// Pi = A[P,P]^-1 and
//     C  = A[:,P] - I[:,P],        Rt = Pi (A[P,:] + I[P,:]),
// the single rank-4 update A <- A - C Rt is exactly the in-place Gauss-Jordan
// step of the block: A[P,P] -> Pi, A[P,j] -> Pi A[P,j], A[i,P] -> -A[i,P] Pi,
// A[i,j] -> A[i,j] - A[i,P] Pi A[P,j].  (Row block P: A_Pj - (A_PP - I) Pi A_Pj
// = Pi A_Pj; column block: A_iP - A_iP (I + Pi) = -A_iP Pi; pivot block:
// A_PP - (A_PP - I)(I + Pi) = Pi.)  On the matrix cores that is four real
// 16x16x4 MFMAs per complex tile and block.  After ceil(n/4) blocks A holds
// A^-1, whose T x T block (T-last order) is the pole's Green's function on T.
// No pivoting: Im A = Im(z) S_W is positive definite for Im z > 0, so every
// leading block (and every Schur complement block) is nonsingular.
//
// Symmetry: A is complex symmetric, so after the pivots of a set D the working
// matrix is [[A_DD^-1, A_DD^-1 A_DU], [-A_UD A_DD^-1, A_UU - A_UD A_DD^-1 A_DU]]
// (U the rest): symmetric on D x D and U x U, antisymmetric between them
// (W[u][d] = -W[d][u]).  Only the tiles on and above the diagonal are needed:
// the column panel A[:,P] is the row panel A[P,:] with the sign flipped on the
// rows already pivoted (i < 4p), and the row panel left of the pivot tile is
// the (negated, transposed) tile column above it.  Wave w owns the tile columns
// J = w, w + NWC, ...; its k-th column keeps the tile rows I < NWC (k + 1) (all
// of the upper tiles of that column, plus at most NWC - 1 below the diagonal,
// which are updated like the others and never read).  The upper tiles see the
// same arithmetic for every NWC, so every split gives the same bits.
//
// Per block the owners of the pivot tile row (and, left of the pivot tile, of
// the pivot tile column) publish A[P,:] to LDS (double buffered: one barrier
// per block; real and imaginary parts in separate arrays).  Every wave then
// inverts the 4 x 4 pivot block on 16 lanes (lane 4i + j holds entry (i, j);
// the pivot row, column and entry come from ds_swizzle), forms its operands C
// and Rt from the panel and updates its tiles.
//
// Accumulator layouts (checked on gfx950 with one-hot operands):
//   v_mfma_f32_16x16x4f32: lane l, register r -> row 4 (l / 16) + r, column l % 16
//   v_mfma_f64_16x16x4f64: lane l, register r -> row l / 16 + 4 r,   column l % 16
//   both: A operand lane l -> A[l % 16][l / 16] (16 x 4), B operand lane l -> B[l / 16][l % 16] (4 x 16).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "lres_device.h"
#include "lres_internal.h"

namespace lres {
namespace {

template <typename R>
struct Mfma;

template <>
struct Mfma<float> {
    typedef float V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

template <>
struct Mfma<double> {
    typedef double V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};

// ds_swizzle in bit-mask mode: source lane ((lane & AND) | OR) within each 32-lane half
template <int PAT>
__device__ __forceinline__ float swz(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), PAT));
}

template <int PAT>
__device__ __forceinline__ double swz(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)b, PAT);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), PAT);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int PAT, typename R>
__device__ __forceinline__ cx<R> swzc(cx<R> v) {
    return {swz<PAT>(v.re), swz<PAT>(v.im)};
}

// Every product-sum below is written with explicit fma and this file is compiled
// with -ffp-contract=off: no rounding is left to the compiler's choice, so all
// waves-per-matrix variants give bitwise the same results.

// 1 / a: the hardware reciprocal (1 ulp) in fp32, a correctly rounded division in fp64
__device__ __forceinline__ cx<float> crecip_mf(cx<float> a) {
    const float s = __builtin_amdgcn_rcpf(fmaf(a.re, a.re, a.im * a.im));
    return {a.re * s, -a.im * s};
}

__device__ __forceinline__ cx<double> crecip_mf(cx<double> a) {
    const double s = 1.0 / fma(a.re, a.re, a.im * a.im);
    return {a.re * s, -a.im * s};
}

template <typename R>
__device__ __forceinline__ cx<R> cmul_mf(cx<R> a, cx<R> b) {
    return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}

// One in-place Gauss-Jordan pivot S of the 4 x 4 block held as lane 4i + j ->
// entry (i, j) in every 16-lane group, with the same shifted form as the block
// update: a_ij -= (a_iS - d_iS)(a_Sj + d_Sj) / a_SS.
template <int S, typename R>
__device__ __forceinline__ void pivot4(cx<R>& a, int i, int j) {
    constexpr int kRowS = 0x1C | (S << 5);           // lane 4i + S (keep bits 4..2, or S)
    constexpr int kColS = 0x13 | ((4 * S) << 5);     // lane 4S + j (keep bits 4, 1..0, or 4S)
    constexpr int kPiv = 0x10 | ((5 * S) << 5);      // lane 5S     (keep bit 4, or 5S)
    cx<R> ci = swzc<kRowS>(a);
    cx<R> rj = swzc<kColS>(a);
    const cx<R> piv = swzc<kPiv>(a);
    if (i == S) ci.re -= R(1);
    if (j == S) rj.re += R(1);
    const cx<R> rs = cmul_mf(rj, crecip_mf(piv));
    a.re = fma(-ci.re, rs.re, fma(ci.im, rs.im, a.re));
    a.im = fma(-ci.re, rs.im, fma(-ci.im, rs.re, a.im));
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// NP > 1 (complex64, NWC = 1): the window-fused form.  One workgroup of NP waves per
// window, wave q sweeping poles q, q + NP, ... with a row panel of its own in LDS (one
// wave per matrix: its LDS operations complete in order, so one buffer and a wave-level
// wait replace the double buffer and the workgroup barrier).  After each round of NP
// poles the waves add their Re(w G) and Re(w z G), rounded to float as the per-pole
// slots hold them, to double accumulators in LDS in wave (= pole) order; at the end the
// workgroup forms D_B, F_B, N_B and E_B exactly as k_reduce does from the slots (the
// same additions in the same order, the trace sums over k_reduce's 1024 virtual threads
// and reduction tree), so both forms give the same bits.
// Tile columns of the matrix-core pole kernel's waves (see k_pole_mfma): with NWC > 1,
// round j deals columns NT - 1 - j NWC down to NT - NWC - j NWC to waves 0..NWC-1, in
// reverse wave order on odd rounds.  held: tile rows kept for a wave's j-th column (the
// largest of the round's columns + 1); owner and index: where tile column c lives.
template <int NT, int NWC>
__host__ __device__ constexpr int pole_held(int j) {
    if (NWC == 1) return j + 1 < NT ? j + 1 : NT;
    return NT - j * NWC;
}
template <int NT, int NWC>
__host__ __device__ constexpr int pole_col_index(int c) {
    return NWC == 1 ? c : (NT - 1 - c) / NWC;
}
template <int NT, int NWC>
__host__ __device__ constexpr int pole_col_owner(int c) {
    if (NWC == 1) return 0;
    const int j = (NT - 1 - c) / NWC, q = (NT - 1 - c) % NWC;
    return (j & 1) ? NWC - 1 - q : q;
}
// the column a wave holds as its j-th (the kernel's cmap), -1 for none
template <int NT, int NWC>
__host__ __device__ constexpr int pole_col_of(int wv, int j) {
    if (NWC == 1) return j;
    return NT - 1 - (j * NWC + ((j & 1) ? NWC - 1 - wv : wv));
}
// every tile column lives in exactly one (wave, j) with its c + 1 upper tiles held, and
// every held (wave, j) names a column or none
template <int NT, int NWC>
constexpr bool pole_map_ok() {
    constexpr int TJ = (NT + NWC - 1) / NWC;
    for (int c = 0; c < NT; ++c) {
        const int w = pole_col_owner<NT, NWC>(c), j = pole_col_index<NT, NWC>(c);
        if (w < 0 || w >= NWC || j < 0 || j >= TJ) return false;
        if (pole_col_of<NT, NWC>(w, j) != c || pole_held<NT, NWC>(j) < c + 1) return false;
    }
    int seen = 0;
    for (int w = 0; w < NWC; ++w)
        for (int j = 0; j < TJ; ++j) {
            const int c = pole_col_of<NT, NWC>(w, j);
            if (c >= NT || c < -NWC) return false;
            if (c >= 0) ++seen;
        }
    return seen == NT;
}

template <typename R, int NT, int NWC, int WPE, int NP>
__global__ void __launch_bounds__(64 * NWC * NP) __attribute__((amdgpu_waves_per_eu(WPE))) k_pole_mfma(const WinDesc* __restrict__ wins,
                                                        const int32_t* __restrict__ ids,
                                                        const double* __restrict__ SW,
                                                        const double* __restrict__ HW, int nids, int M, int ngroup,
                                                        const double* __restrict__ circ, double* __restrict__ Dpart,
                                                        double* __restrict__ Fpart, int acc, WinOut wout) {
    using MF = Mfma<R>;
    using V = typename MF::V;
    static_assert(NP == 1 || (NWC == 1 && sizeof(R) == 4), "the window-fused form is complex64, one wave per matrix");
    constexpr int TJ = (NT + NWC - 1) / NWC;   // tile columns per wave
    constexpr int NR = 16 * NT;                 // padded order
    // row panel of the current block, planar (real and imaginary parts apart) so
    // that accumulator registers go to LDS and come back as MFMA operands without
    // repacking: rp[b][j][s] = A[4p + s][j].  fp64 rows are padded to 40 bytes: the
    // B operands read one column per lane, and at a 32-byte stride lanes j and j + 8
    // would share LDS banks
    constexpr int RS = sizeof(R) == 8 ? 5 : 4;
    constexpr int NBUF = NP > 1 ? NP : 2;
    // NP > 1: between rounds the two planes are the two buffers of the stage of the
    // round's sums (one tile of every wave's float pairs, [wave][register][lane], 8 KB:
    // the planes are 128 rows deep then)
    constexpr int NRP = NP > 1 ? 128 : NR;
    static_assert(NP == 1 || (NR <= NRP && NBUF * NRP * RS == NP * 256 * 2), "stage buffer size");
    __shared__ __attribute__((aligned(16))) R rpr[NBUF][NRP][RS];
    __shared__ __attribute__((aligned(16))) R rpi[NBUF][NRP][RS];
    // fp64 with acc > 0: the group's D and F partial sums over its poles, packed upper
    // triangles of acc entries each (dynamic LDS), stored to Dpart / Fpart once at the
    // end instead of a read-modify-write of global memory per pole; NP > 1: the window's
    // D and F sums over all poles
    extern __shared__ double pacc[];
    // the circle (c, R) of the pass; circ[2] != 0 marks a pass the device-side mu search
    // did not need (every workgroup leaves at once)
    if (circ[2] != 0.0) return;
    const double c = circ[0], Rr = circ[1];
    // XCD-aware order: workgroups are dealt to the 8 XCDs round-robin by linear id, so
    // id L -> window slot 8 (L / (8 ngroup)) + L % 8, group (L / 8) % ngroup puts every
    // group of a window on one XCD, dispatched together: the window's S_W and H_W are
    // then read from HBM about once per pass and from that XCD's L2 by its other groups
    const int L = blockIdx.x;
    const int slot = NP > 1 ? L : (L / (8 * ngroup)) * 8 + (L & 7);
    const int gq = NP > 1 ? 0 : (L >> 3) % ngroup;
    if (slot >= nids) return;
    const int w = ids[slot];
    const WinDesc d = wins[w];
    const int n = d.n, m = d.m, k = d.k;
    const int lane = threadIdx.x & 63;
    // NP > 1: wave `mat` holds its own matrix (pole), every wave is the matrix's wave 0
    const int mat = NP > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    const int wv = NP > 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // the panel buffer of block parity b: NP > 1 one buffer per wave
    auto bsel = [&](int b) __attribute__((always_inline)) { return NP > 1 ? mat : b; };
    // the panel written by some lanes is read by others: a workgroup barrier, or for a
    // wave alone with its panel, the completion of its LDS writes
    auto panel_sync = [&]() __attribute__((always_inline)) {
        if constexpr (NP > 1) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0)
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
    };
    const double* S = SW + d.mat_off;
    const double* H = HW + d.mat_off;
    // fp64: this group's partial sums (Dpart, Fpart); fp32: Dpart is the per-pole
    // slot array, (M / 2) k x k float pairs per window
    double* dp = Dpart + d.kk_off * ngroup + (int64_t)gq * k * k;
    double* fp = Fpart + d.kk_off * ngroup + (int64_t)gq * k * k;
    float2* gs = reinterpret_cast<float2*>(Dpart) + d.kk_off * (M / 2);
    const int nst = (n + 3) >> 2;

    // are[I][j], aim[I][j]: tile (I, cmap(j)), held for I < held(j) only.  One wave per
    // matrix holds tile column j as its j-th; with NWC > 1 the columns (column c has c + 1
    // upper tiles) are dealt largest first in snake order, so that the waves' shares of
    // the upper triangle balance (NT 6: {5, 0} {4, 1} {3, 2} over 3 waves, {5} {4} {3, 0}
    // {2, 1} over 4).  Every wave keeps and updates held(j) tile rows of its j-th column
    // (the same register layout in every wave; the tiles below the diagonal are never
    // read), so a wave's updates per block step are the sum of held(j): 8 tiles at NT 6 over
    // 4 waves (10 with columns dealt round-robin), 7 over 5 (11)
    V are[NT][TJ], aim[NT][TJ];
    static_assert(pole_map_ok<NT, NWC>(), "tile columns must map one to one onto (wave, j)");
    auto cmap = [&](int j) __attribute__((always_inline)) { return pole_col_of<NT, NWC>(wv, j); };
    auto owned = [&](int j) __attribute__((always_inline)) { return NWC == 1 || cmap(j) >= 0; };
    auto held = [](int j) __attribute__((always_inline)) { return pole_held<NT, NWC>(j); };

    // pivot rows of block p = 4 JPc + pq into buffer b: tile row JPc of the owned
    // columns J >= JPc, and left of it the tile column JPc above the diagonal,
    // transposed.  Left of JPc the row panel is the negated transpose (those rows are
    // pivoted, the pivot rows are not): it is stored without the sign, which the
    // consumers fold in (the A operands of the pivoted rows would negate it back; the
    // B operands of columns J < JP take the opposite signs of their products)
    auto publish = [&](auto jpc, int pq, int b, int li, int lg) __attribute__((always_inline)) {
        constexpr int JPc = decltype(jpc)::value;
        constexpr int jc = pole_col_index<NT, NWC>(JPc);
        if (JPc > 0 && wv == pole_col_owner<NT, NWC>(JPc) && (li >> 2) == pq) {
            const int sc = li & 3;
#pragma unroll
            for (int I = 0; I < JPc; ++I)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    rpr[bsel(b)][16 * I + MF::row(lg, r)][sc] = are[I][jc][r];
                    rpi[bsel(b)][16 * I + MF::row(lg, r)][sc] = aim[I][jc][r];
                }
        }
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
            if (JPc >= held(j) || !owned(j)) continue;
            if (cmap(j) < JPc) continue;
            const int col = 16 * cmap(j) + li;
            if constexpr (sizeof(R) == 4) {   // lane group pq holds rows 4 pq + s in register s
                if (lg == pq) {
                    *reinterpret_cast<V*>(&rpr[bsel(b)][col][0]) = are[JPc][j];
                    *reinterpret_cast<V*>(&rpi[bsel(b)][col][0]) = aim[JPc][j];
                }
            } else {                          // register pq holds row 4 pq + lg
                const V& xr = are[JPc][j];
                const V& xi = aim[JPc][j];
                rpr[bsel(b)][col][lg] = pq == 0 ? xr[0] : pq == 1 ? xr[1] : pq == 2 ? xr[2] : xr[3];
                rpi[bsel(b)][col][lg] = pq == 0 ? xi[0] : pq == 1 ? xi[1] : pq == 2 ? xi[2] : xi[3];
            }
        }
        // distinct per tile: stops the two publish branches of a block from being
        // tail-merged into one with a run-time tile index (which would move the
        // accumulators to scratch memory)
        asm volatile("; publish tile %0" ::"n"(JPc));
    };

    bool first = true;
    const int npole = M / 2;
    // NP > 1: rounds of NP poles (wave mat takes pole NP r + mat; a wave without a pole in
    // the last round still takes part in its barriers); else the group's poles in turn
    const int nround = NP > 1 ? (npole + NP - 1) / NP : (npole - gq + ngroup - 1) / ngroup;
    for (int rd = 0; rd < nround; ++rd) {
        const int jp = NP > 1 ? NP * rd + mat : gq + ngroup * rd;
        const bool active = NP == 1 || jp < npole;
        cx<double> z, wt, wz;
        pole(active ? jp : 0, M, c, Rr, z, wt, wz);
        if (active) {
        // the order made opaque per pole: keeps the per-entry offsets and bounds
        // tests of the load from being held in registers across the sweep
        int nn = n, ln0 = lane;
        asm volatile("" : "+s"(nn), "+v"(ln0));
        const int li = ln0 & 15, lg = ln0 >> 4;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                if (I >= held(j) || !owned(j)) continue;   // an unowned column is never read
                const int col = 16 * cmap(j) + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * I + MF::row(lg, r);
                    // every entry loaded (entry 0 stands in for the padding) and
                    // masked arithmetically, so no branch per entry; the padding is
                    // 0 off and 1 on the diagonal
                    const bool in = row < nn && col < nn;
                    // 32-bit element offset (n <= 128): base-register addressing, one VGPR per address
                    const uint32_t off = in ? (uint32_t)(row * nn + col) : 0u;
                    const double f = in ? 1.0 : 0.0;
                    const double pad = (!in && row == col) ? 1.0 : 0.0;
                    const double s = S[off] * f;
                    const double h = H[off] * f;
                    are[I][j][r] = (R)(fma(z.re, s, pad) - h);
                    aim[I][j][r] = (R)(z.im * s);
                }
            }
            // this tile row's entries are formed before the next row's loads issue
            // (the memory clobber keeps those loads below this point)
#pragma unroll
            for (int j = 0; j < TJ; ++j)
                if (I < held(j)) asm volatile("" : "+v"(are[I][j]), "+v"(aim[I][j])::"memory");
        }
        {
            int ln = lane;
            asm volatile("" : "+v"(ln));
            publish(std::integral_constant<int, 0>{}, 0, 0, ln & 15, ln >> 4);
        }

        static_for<0, NT>([&](auto jpc) __attribute__((always_inline)) {
            constexpr int JP = decltype(jpc)::value;
            const int ppend = min(4, nst - 4 * JP);
            for (int pp = 0; pp < ppend; ++pp) {
                const int p = 4 * JP + pp, b = p & 1;
                panel_sync();
                // lane indices made opaque per block: per-lane LDS addresses and
                // masks are recomputed here rather than held across the sweep
                int ln = lane;
                asm volatile("" : "+v"(ln));
                const int li = ln & 15, lg = ln >> 4;
                // Pi = A[P,P]^-1 on 16 lanes (four identical copies per wave)
                const int pi = li >> 2, pj = li & 3;
                cx<R> a = {rpr[bsel(b)][4 * p + pj][pi], rpi[bsel(b)][4 * p + pj][pi]};
                pivot4<0>(a, pi, pj);
                pivot4<1>(a, pi, pj);
                pivot4<2>(a, pi, pj);
                pivot4<3>(a, pi, pj);
                cx<R> prow[4];   // row lg of Pi
#pragma unroll
                for (int s = 0; s < 4; ++s) prow[s] = {__shfl(a.re, 4 * lg + s, 64), __shfl(a.im, 4 * lg + s, 64)};
                // A operands: C[16 I + li][lg] = A[16 I + li][4p + lg] - d, read from the
                // row panel with the sign flipped on the pivoted rows (16 I + li < 4p; the
                // tiles I < JP are stored without the panel's sign, so only the pivoted
                // rows of tile row JP flip here); the identity shift only in tile row JP
                R car[NT], cai[NT];
#pragma unroll
                for (int I = 0; I < NT; ++I) {
                    R xr = rpr[bsel(b)][16 * I + li][lg];
                    R xi = rpi[bsel(b)][16 * I + li][lg];
                    if (I == JP && li < 4 * pp) {
                        xr = -xr;
                        xi = -xi;
                    }
                    if (I == JP && li == 4 * pp + lg) xr -= R(1);
                    car[I] = xr;
                    cai[I] = xi;
                }
                // B operands: Rt[lg][16 J + li] = sum_s Pi[lg][s] (A[4p + s][col] + d), the
                // identity shift only in tile column JP
                R nre[TJ], rim[TJ], nim[TJ];
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    nre[j] = rim[j] = nim[j] = R(0);
                    if (!owned(j)) continue;
                    const int col = 16 * cmap(j) + li;
                    R tr[4], ti[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        tr[s] = rpr[bsel(b)][col][s];
                        ti[s] = rpi[bsel(b)][col][s];
                    }
                    if (cmap(j) == JP) {
#pragma unroll
                        for (int s = 0; s < 4; ++s)
                            if (li == 4 * pp + s) tr[s] += R(1);
                    }
                    R xr = R(0), xi = R(0);
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        xr = fma(prow[s].re, tr[s], fma(-prow[s].im, ti[s], xr));
                        xi = fma(prow[s].re, ti[s], fma(prow[s].im, tr[s], xi));
                    }
                    // columns left of the pivot tile read the panel without its sign:
                    // their products come out negated (exactly), so the signs swap
                    const bool left = cmap(j) < JP;
                    nre[j] = left ? xr : -xr;
                    rim[j] = left ? -xi : xi;
                    nim[j] = left ? xi : -xi;
                }
                // A <- A - C Rt:  Re -= Cr Rr - Ci Ri,  Im -= Cr Ri + Ci Rr
#pragma unroll
                for (int I = 0; I < NT; ++I)
#pragma unroll
                    for (int j = 0; j < TJ; ++j) {
                        if (I >= held(j) || !owned(j)) continue;
                        are[I][j] = MF::mma(car[I], nre[j], are[I][j]);
                        are[I][j] = MF::mma(cai[I], rim[j], are[I][j]);
                        aim[I][j] = MF::mma(car[I], nim[j], aim[I][j]);
                        aim[I][j] = MF::mma(cai[I], nre[j], aim[I][j]);
                    }
                if (p + 1 < nst) {
                    if (pp < 3)
                        publish(jpc, pp + 1, b ^ 1, li, lg);
                    else
                        publish(std::integral_constant<int, (JP + 1 < NT ? JP + 1 : JP)>{}, 0, b ^ 1, li, lg);
                }
            }
        });
        panel_sync();   // the next pole's first publish rewrites buffer 0
        }   // active

        if constexpr (NP > 1) {
            // this round's poles into the window's sums, one tile at a time: every wave
            // stages its Re(w G) and Re(w z G) of the tile, rounded to float as a per-pole
            // slot holds them; then thread (wave q, lane l) adds entry (register q, lane l)
            // of the round's poles, in pole order, to the double sums.  The same additions
            // in the same order as the slots and k_reduce, with the four waves adding at
            // once.  The stage buffers are the two panel planes, free once every wave is
            // past its sweep; two buffers, so one barrier per tile.  Tiles with no T row or
            // column are skipped (m is the workgroup's)
            const int ei = lane & 15, eg = lane >> 4;
            int tb = 0;
            __syncthreads();
#pragma unroll
            for (int I = 0; I < NT; ++I)
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    if (I >= held(j) || 16 * (j + 1) <= m || 16 * (I + 1) <= m) continue;
                    float* sb = (tb & 1) ? &rpi[0][0][0] : &rpr[0][0][0];
                    ++tb;
                    if (active) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const double gr = (double)are[I][j][r], gi = (double)aim[I][j][r];
                            *reinterpret_cast<float2*>(sb + 2 * ((4 * mat + r) * 64 + lane)) =
                                make_float2((float)fma(wt.re, gr, -(wt.im * gi)), (float)fma(wz.re, gr, -(wz.im * gi)));
                        }
                    }
                    __syncthreads();
                    const int col = 16 * j + ei - m;
                    const int row = 16 * I + MF::row(eg, mat) - m;
                    if (row >= 0 && row <= col && col < k) {
                        const int u = row * k - ((row * (row - 1)) >> 1) + (col - row);
                        double sd = rd == 0 ? 0.0 : pacc[u];
                        double sf = rd == 0 ? 0.0 : pacc[acc + u];
                        for (int q = 0; q < NP && NP * rd + q < npole; ++q) {
                            const float2 g = *reinterpret_cast<const float2*>(sb + 2 * ((4 * q + mat) * 64 + lane));
                            sd += (double)g.x;
                            sf += (double)g.y;
                        }
                        pacc[u] = sd;
                        pacc[acc + u] = sf;
                    }
                }
            __syncthreads();   // the stage is the next round's panel memory
            continue;
        }

        // Re(w G) and Re(w z G) on T x T: fp64 adds them to this group's partial
        // sums; fp32 stores them once, as a float pair, in this pole's own slot (no
        // read-modify-write; the reduction sums the poles in order)
        int kk = k, mm = m, ln1 = lane;
        asm volatile("" : "+s"(kk), "+s"(mm), "+v"(ln1));
        const int ei = ln1 & 15, eg = ln1 >> 4;
#pragma unroll
        for (int I = 0; I < NT; ++I)
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                if (I >= held(j) || !owned(j)) continue;
                const int col = 16 * cmap(j) + ei - mm;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * I + MF::row(eg, r) - mm;
                    if (row >= 0 && row <= col && col < kk) {   // upper triangle: G[T,T] is symmetric
                        const double gr = (double)are[I][j][r], gi = (double)aim[I][j][r];
                        const double dv = fma(wt.re, gr, -(wt.im * gi)), fv = fma(wz.re, gr, -(wz.im * gi));
                        const int64_t q = (int64_t)row * kk + col;
                        if constexpr (sizeof(R) == 4) {
                            gs[(int64_t)jp * kk * kk + q] = make_float2((float)dv, (float)fv);
                        } else if (acc > 0) {
                            const int u = row * kk - ((row * (row - 1)) >> 1) + (col - row);
                            if (first) {
                                pacc[u] = dv;
                                pacc[acc + u] = fv;
                            } else {
                                pacc[u] += dv;
                                pacc[acc + u] += fv;
                            }
                        } else if (first) {
                            dp[q] = dv;
                            fp[q] = fv;
                        } else {
                            dp[q] += dv;
                            fp[q] += fv;
                        }
                    }
                }
            }
        first = false;
    }
    if constexpr (NP > 1) {
        // D_B, F_B (upper triangles), N_B, E_B as k_reduce forms them from the slots: the
        // entries dealt to 1024 virtual threads (thread t of this workgroup plays t + 64 NP v,
        // v < 1024 / (64 NP)), fma trace sums, each virtual wave's shuffle tree, then the 16
        // wave partials in order
        static_assert(NP == 4, "the virtual-thread mapping assumes four waves");
        __shared__ double red[2][16];
        __syncthreads();
        const int nup = k * (k + 1) / 2;
        auto row_start = [k](int r) { return r * k - (r * (r - 1)) / 2; };
        double* Dw = wout.D + d.kk_off;
        double* Fw = wout.F + d.kk_off;
        const double* Qw = wout.Q + d.kk_off;
#pragma unroll
        for (int vq = 0; vq < 4; ++vq) {
            double sn = 0.0, se = 0.0;
            for (int e = threadIdx.x + 256 * vq; e < nup; e += 1024) {
                const float b2 = 2.0f * (float)k + 1.0f;
                int i = (int)(0.5f * (b2 - sqrtf(fmaxf(b2 * b2 - 8.0f * (float)e, 0.0f))));
                i = min(max(i, 0), k - 1);
                while (i > 0 && row_start(i) > e) --i;
                while (i + 1 < k && row_start(i + 1) <= e) ++i;
                const int j = i + (e - row_start(i));
                const int64_t q = (int64_t)i * k + j;
                const double dv = pacc[e] * 2.0, fv = pacc[acc + e] * 2.0;
                Dw[q] = dv;
                Fw[q] = fv;
                double qv = Qw[q];
                if (j != i) qv *= 2.0;
                sn = fma(dv, qv, sn);
                se = fma(fv, qv, se);
            }
            for (int o = 32; o > 0; o >>= 1) {
                sn += __shfl_down(sn, o, 64);
                se += __shfl_down(se, o, 64);
            }
            if (lane == 0) {
                red[0][mat + 4 * vq] = sn;
                red[1][mat + 4 * vq] = se;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tn = 0.0, te = 0.0;
            for (int q = 0; q < 16; ++q) tn += red[0][q];
            for (int q = 0; q < 16; ++q) te += red[1][q];
            if (wout.NB) wout.NB[w] = tn;
            if (wout.EB) wout.EB[w] = te;
        }
        return;
    }
    if constexpr (sizeof(R) == 8) {
        // each lane stores the entries it accumulated (no barrier: the same lane wrote them)
        if (acc > 0) {
            const int ei = lane & 15, eg = lane >> 4;
#pragma unroll
            for (int I = 0; I < NT; ++I)
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    if (I >= held(j) || !owned(j)) continue;
                    const int col = 16 * cmap(j) + ei - m;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * I + MF::row(eg, r) - m;
                        if (row >= 0 && row <= col && col < k) {
                            const int u = row * k - ((row * (row - 1)) >> 1) + (col - row);
                            const int64_t q = (int64_t)row * k + col;
                            dp[q] = pacc[u];
                            fp[q] = pacc[acc + u];
                        }
                    }
                }
        }
    }
}

// One real Gauss-Jordan pivot S of the 4 x 4 block held as lane 4i + j -> entry
// (i, j) (the shifted form of pivot4); returns the pivot value, which is the
// S-th pivot of the block's unpivoted LDL^T.
template <int S>
__device__ __forceinline__ double rpivot4(double& a, int i, int j) {
    constexpr int kRowS = 0x1C | (S << 5);
    constexpr int kColS = 0x13 | ((4 * S) << 5);
    constexpr int kPiv = 0x10 | ((5 * S) << 5);
    double ci = swz<kRowS>(a);
    double rj = swz<kColS>(a);
    const double piv = swz<kPiv>(a);
    if (i == S) ci -= 1.0;
    if (j == S) rj += 1.0;
    a = fma(-ci, rj / piv, a);
    return piv;
}

// Inertia of sigma S_W - H_W (number of positive pivots of its unpivoted LDL^T,
// the same count as the vector-ALU k_inertia) on the matrix cores: block Gaussian
// elimination in blocks of four pivots, one wave per (window, shift).  Only the
// trailing Schur complement matters and it stays symmetric, so the wave holds the
// tiles on and above the diagonal, updates at block p only those at or below the
// pivot tile row (one v_mfma_f64_16x16x4f64 each: A -= A[:,P] (A_PP^-1 A[P,:])),
// and takes the column panel from the row panel.  The four pivots of a block come
// out of the 4 x 4 inverse of its diagonal block (rpivot4, in order); pivots on the
// identity padding are not counted.
template <int NT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NT <= 6 ? 2 : 1)))
k_inertia_mfma(const WinDesc* __restrict__ wins, const int32_t* __restrict__ ids, const double* __restrict__ SW,
               const double* __restrict__ HW, const double* __restrict__ sigma, int nshift,
               int32_t* __restrict__ counts, unsigned long long* __restrict__ total, const int* __restrict__ skip,
               const int* __restrict__ lin) {
    using MF = Mfma<double>;
    using V = typename MF::V;
    constexpr int NR = 16 * NT;
    // row panel of the current block: rp[b][j][s] = A[4p + s][j], rows padded to 40
    // bytes (a 32-byte stride puts the column reads of lanes j and j + 8 on the same
    // LDS banks)
    __shared__ __attribute__((aligned(16))) double rp[2][NR][5];
    if (skip != nullptr && *skip != 0) return;   // decided by the keep-edge bound
    int c = blockIdx.x, s_id = blockIdx.y;
    if (lin != nullptr) {   // linear launch of a device lambda_min round: (candidate, shift)
        if (lin[2] != 0) return;
        const int K = lin[1];
        c = blockIdx.x / K;
        if (c >= lin[0]) return;
        s_id = blockIdx.x - c * K;
    }
    const int w = ids[c];
    const WinDesc d = wins[w];
    const int n = d.n;
    const int lane = threadIdx.x & 63;
    const double sg = sigma[s_id];
    const double* S = SW + d.mat_off;
    const double* H = HW + d.mat_off;
    const int nst = (n + 3) >> 2;

    V a[NT][NT];   // tile (I, J) for I <= J only
    {
        int nn = n, ln0 = lane;
        asm volatile("" : "+s"(nn), "+v"(ln0));
        const int li = ln0 & 15, lg = ln0 >> 4;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int J = 0; J < NT; ++J) {
                if (J < I) continue;
                const int col = 16 * J + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * I + MF::row(lg, r);
                    const bool in = row < nn && col < nn;
                    const uint32_t off = in ? (uint32_t)(row * nn + col) : 0u;
                    const double sv = S[off], hv = H[off];
                    a[I][J][r] = in ? sg * sv - hv : (row == col ? 1.0 : 0.0);
                }
            }
#pragma unroll
            for (int J = 0; J < NT; ++J)
                if (J >= I) asm volatile("" : "+v"(a[I][J])::"memory");
        }
    }
    // tile row JPc (columns of tiles J >= JPc) of block 4 JPc + pq into buffer b;
    // register pq of lane group s holds row 4 pq + s
    auto publish = [&](auto jpc, int pq, int b, int li, int lg) __attribute__((always_inline)) {
        constexpr int JPc = decltype(jpc)::value;
#pragma unroll
        for (int J = 0; J < NT; ++J) {
            if (J < JPc) continue;
            const V& x = a[JPc][J];
            rp[b][16 * J + li][lg] = pq == 0 ? x[0] : pq == 1 ? x[1] : pq == 2 ? x[2] : x[3];
        }
        asm volatile("; inertia publish tile %0" ::"n"(JPc));
    };
    {
        int ln = lane;
        asm volatile("" : "+v"(ln));
        publish(std::integral_constant<int, 0>{}, 0, 0, ln & 15, ln >> 4);
    }
    int npos = 0;
    static_for<0, NT>([&](auto jpc) __attribute__((always_inline)) {
        constexpr int JP = decltype(jpc)::value;
        const int ppend = min(4, nst - 4 * JP);
        for (int pp = 0; pp < ppend; ++pp) {
            const int p = 4 * JP + pp, b = p & 1;
            __syncthreads();
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int li = ln & 15, lg = ln >> 4;
            const int pi = li >> 2, pj = li & 3;
            double x = rp[b][4 * p + pj][pi];
            const double v0 = rpivot4<0>(x, pi, pj);
            const double v1 = rpivot4<1>(x, pi, pj);
            const double v2 = rpivot4<2>(x, pi, pj);
            const double v3 = rpivot4<3>(x, pi, pj);
            npos += (v0 > 0.0 && 4 * p < n ? 1 : 0) + (v1 > 0.0 && 4 * p + 1 < n ? 1 : 0) +
                    (v2 > 0.0 && 4 * p + 2 < n ? 1 : 0) + (v3 > 0.0 && 4 * p + 3 < n ? 1 : 0);
            double prow[4];   // row lg of A_PP^-1
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2) prow[s2] = __shfl(x, 4 * lg + s2, 64);
            // A operands: C[16 I + li][lg] = A[4p + lg][16 I + li] on the trailing rows, 0 on
            // the rows already eliminated and the pivot rows
            double car[NT];
#pragma unroll
            for (int I = 0; I < NT; ++I) {
                car[I] = 0.0;
                if (I < JP) continue;
                const int row = 16 * I + li;
                const double v = rp[b][row][lg];
                car[I] = row >= 4 * p + 4 ? v : 0.0;
            }
            // B operands: -(A_PP^-1 A[P,:])[lg][16 J + li]
            double nb[NT];
#pragma unroll
            for (int J = 0; J < NT; ++J) {
                nb[J] = 0.0;
                if (J < JP) continue;
                const int col = 16 * J + li;
                double t = 0.0;
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) t = fma(prow[s2], rp[b][col][s2], t);
                nb[J] = -t;
            }
#pragma unroll
            for (int I = 0; I < NT; ++I)
#pragma unroll
                for (int J = 0; J < NT; ++J) {
                    if (I < JP || J < I) continue;
                    a[I][J] = MF::mma(car[I], nb[J], a[I][J]);
                }
            if (p + 1 < nst) {
                if (pp < 3)
                    publish(jpc, pp + 1, b ^ 1, li, lg);
                else
                    publish(std::integral_constant<int, (JP + 1 < NT ? JP + 1 : JP)>{}, 0, b ^ 1, li, lg);
            }
        }
    });
    if (lane == 0) {
        if (counts) counts[(int64_t)w * nshift + s_id] = npos;
        atomicAdd(&total[s_id], (unsigned long long)npos);
    }
}

// Accumulator registers per lane of a variant: 8 (fp32) or 16 (fp64) per held tile.
template <typename R, int NT, int NWC>
constexpr int mfma_acc() {
    constexpr int TJ = (NT + NWC - 1) / NWC;
    int t = 0;
    for (int j = 0; j < TJ; ++j) t += pole_held<NT, NWC>(j);
    return t * (sizeof(R) == 4 ? 8 : 16);
}

// Waves per SIMD the register allocation is held to: as many as leave about 70
// registers beside the accumulators for operands, addresses and the load prologue.
template <typename R, int NT, int NWC>
constexpr int mfma_waves() {
    constexpr int acc = mfma_acc<R, NT, NWC>();
    return acc <= 48 ? 4 : acc <= 80 ? 3 : acc <= 168 ? 2 : 1;
}

template <typename R, int NT, int NWC, int NP = 1>
hipError_t launch_one(const WinDesc* wins, const int32_t* ids, int nids, const double* SW, const double* HW, int M,
                      int ngroup, const double* circ, double* Dpart, double* Fpart, int acc, WinOut out,
                      hipStream_t st) {
    const auto kern = k_pole_mfma<R, NT, NWC, mfma_waves<R, NT, NWC>(), NP>;
    // dynamic LDS: the fp64 group sums or the window-fused form's window sums (acc
    // entries each of D and F)
    const size_t lds = (NP > 1 || sizeof(R) == 8) && acc > 0 ? (size_t)2 * acc * sizeof(double) : 0;
    if (NP > 1 && acc <= 0) return hipErrorInvalidValue;
    if (lds > 0) {
        static bool raised = false;   // the dynamic LDS limit, raised once per variant
        if (!raised) {
            const hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds > 65536 ? (int)lds : 65536);
            if (e != hipSuccess) return e;
            raised = true;
        }
    }
    // one id per (window slot, group), slots past nids exit; the fused form: one per window
    const int nblk = NP > 1 ? nids : (nids + 7) / 8 * 8 * ngroup;
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(64 * NWC * NP), lds, st, wins, ids, SW, HW, nids, M, ngroup, circ, Dpart,
                       Fpart, lds > 0 ? acc : 0, out);
    return hipGetLastError();
}

// private (scratch) bytes per lane of a variant: nonzero means register spills
template <typename R, int NT, int NWC, int NP = 1>
int scratch_one() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_pole_mfma<R, NT, NWC, mfma_waves<R, NT, NWC>(), NP>)) !=
        hipSuccess)
        return -1;
    return (int)a.localSizeBytes;
}

}  // namespace
}  // namespace lres

// One object per (precision, NT): the build compiles this file once per part
// with LRES_MFMA_R / LRES_MFMA_NT set (in parallel), and the dispatcher in
// lres_kernels.hip calls the parts by name.
#if defined(LRES_MFMA_R) && defined(LRES_MFMA_NT)
#define LRES_MFMA_NAME3(a, b, c) a##b##_##c
#define LRES_MFMA_NAME(r, nt) LRES_MFMA_NAME3(lres_mfma_part_, r, nt)
// up to 8 tile columns: every split of 1..8 waves; beyond (fp32 windows of 129..256
// functions, whose tiles fill the register file of eight waves): 8 waves
#if LRES_MFMA_NT <= 8
#define LRES_MFMA_FOR_NWC(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8)
#else
#define LRES_MFMA_FOR_NWC(X) X(8)
#endif
// nwc > 0: launch with nwc waves per (window, pole group); nwc == 0 (fp32, up to 8 tile
// columns): the window-fused form (four waves per window); nwc < 0: return the scratch
// bytes per lane of the -nwc variant (kFusedQuery: of the window-fused form; -1 if the
// variant does not exist)
hipError_t LRES_MFMA_NAME(LRES_MFMA_R, LRES_MFMA_NT)(int nwc, const lres::WinDesc* wins, const int32_t* ids, int nids,
                                                     const double* SW, const double* HW, int M, int ngroup,
                                                     const double* circ, double* Dpart, double* Fpart, int acc,
                                                     lres::WinOut out, hipStream_t st) {
    using R = LRES_MFMA_R;
    constexpr int NT = LRES_MFMA_NT;
    constexpr int kFusedQuery = -100;
#define LRES_MFMA_CASE(W)                                                                                     \
    if constexpr (W <= NT) {                                                                                  \
        if (nwc == W)                                                                                         \
            return lres::launch_one<R, NT, W>(wins, ids, nids, SW, HW, M, ngroup, circ, Dpart, Fpart, acc, out, \
                                              st);                                                            \
        if (nwc == -W) return (hipError_t)lres::scratch_one<R, NT, W>();                                      \
    }
    LRES_MFMA_FOR_NWC(LRES_MFMA_CASE)
#undef LRES_MFMA_CASE
    if constexpr (sizeof(R) == 4 && NT <= 8) {
        if (nwc == 0)
            return lres::launch_one<R, NT, 1, 4>(wins, ids, nids, SW, HW, M, ngroup, circ, Dpart, Fpart, acc, out, st);
        if (nwc == kFusedQuery) return (hipError_t)lres::scratch_one<R, NT, 1, 4>();
    }
    return nwc < 0 ? (hipError_t)-1 : hipErrorInvalidValue;
}

// the inertia kernel: fp64 parts only
#define LRES_MFMA_CAT_(a, b) a##b
#define LRES_MFMA_CAT(a, b) LRES_MFMA_CAT_(a, b)
#define LRES_MFMA_IS_DOUBLE_double 1
#if LRES_MFMA_CAT(LRES_MFMA_IS_DOUBLE_, LRES_MFMA_R)
hipError_t LRES_MFMA_CAT(lres_mfma_inertia_, LRES_MFMA_NT)(const lres::WinDesc* wins, const int32_t* ids, int nids,
                                                          const double* SW, const double* HW, const double* sigma,
                                                          int nshift, int32_t* counts, unsigned long long* total,
                                                          const int* skip, const int* lin, hipStream_t st) {
    const dim3 grid = lin != nullptr ? dim3(nids, 1) : dim3(nids, nshift);
    hipLaunchKernelGGL((lres::k_inertia_mfma<LRES_MFMA_NT>), grid, dim3(64), 0, st, wins, ids, SW, HW, sigma, nshift,
                       counts, total, skip, lin);
    return hipGetLastError();
}
#endif
#endif
