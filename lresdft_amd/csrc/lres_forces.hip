// Atomic forces of the band energy on the device: the chain rule from the window
// step's gradients (the adjoint's dE/dH_W, dE/dS_W and the pass's D_B, F_B) to the
// atom positions through the basis functions (lrft/lcao3d.py:49-100: Gaussians times
// (1 - (r/rcut)^2)^2, normalised by the grid quadrature).
//
// Every atom's functions live on a dense box of (2 rad + 1)^3 grid points around
// round(R / dx), rad = ceil(rcut / dx), which holds the whole cut-off sphere (the
// box faces lie at >= rcut, where the functions vanish).  The box keeps, per
// function c, x_c, the 7-point kinetic image T x_c (T = -1/2 Laplacian, three_d.py:151-164;
// T x_c vanishes outside the box because x_c vanishes on its faces) and
// dx_c/dR_A (three components).  With G^H, G^S the gradients of E with respect to
// H and S (spin factor included) and G^Q_B that with respect to Q_B:
//   dE/dR_A = 2 dV sum_{c on A} sum_g dx_c(g)/dR_A [ sum_b G^H_cb ((T + v) x_b)(g)
//                                                   + G^S_cb x_b(g) + sum_u G^Q_{B(g),cu} x_u(g) ]
// (H = X^T (T + diag v) X dV, S = X^T X dV, Q_B = sum_{g in own B} x_g x_g^T dV).  The pair
// part runs over atom pairs whose boxes meet, the Q part over each atom's box points.
// Sums have a fixed order (per-thread partials, fixed block trees, per-atom sums over
// a fixed pair list), so results are bitwise reproducible.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lres_internal.h"

namespace lres {
namespace {

constexpr int kFT = 256;   // threads per block

__device__ __forceinline__ double block_sum_f(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    __syncthreads();
    if (ln == 0) red[wv] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    return t;
}

__device__ __forceinline__ int wrap(int64_t i, int n) {
    int64_t r = i % n;
    return (int)(r < 0 ? r + n : r);
}

// Basis functions on the boxes: x (nf per atom), dx/dR (3 nf), T x (nf), each nb3
// values per function, atom-major.  One block per atom.
__global__ void __launch_bounds__(kFT) k_atom_boxes(int natom, const double* __restrict__ pos,
                                                     const int64_t* __restrict__ origin, int nbox, double dx,
                                                     double Lx, double Ly, double Lz, double rcut,
                                                     const double* __restrict__ exps, int ns, int np, double dV,
                                                     double* __restrict__ X, double* __restrict__ DX,
                                                     double* __restrict__ TX) {
    __shared__ double red[kFT / 64];
    __shared__ double nrm[8], dn[8][3];
    const int A = blockIdx.x;
    const int nf = ns + 3 * np;
    const int nb3 = nbox * nbox * nbox;
    const double px = pos[3 * A], py = pos[3 * A + 1], pz = pos[3 * A + 2];
    const int64_t ox = origin[3 * A], oy = origin[3 * A + 1], oz = origin[3 * A + 2];
    double* x = X + (int64_t)A * nf * nb3;
    double* d = DX + (int64_t)A * 3 * nf * nb3;
    double* t = TX + (int64_t)A * nf * nb3;
    const double rc2 = rcut * rcut;
    double s2[8], sd[8][3];
    for (int c = 0; c < 8; ++c) s2[c] = sd[c][0] = sd[c][1] = sd[c][2] = 0.0;
    for (int p = threadIdx.x; p < nb3; p += kFT) {
        const int i = p / (nbox * nbox), j = (p / nbox) % nbox, k = p % nbox;
        double ddx = (double)(ox + i) * dx - px, ddy = (double)(oy + j) * dx - py, ddz = (double)(oz + k) * dx - pz;
        ddx -= Lx * rint(ddx / Lx);
        ddy -= Ly * rint(ddy / Ly);
        ddz -= Lz * rint(ddz / Lz);
        const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
        const bool keep = r2 < rc2;
        const double u = r2 / rc2, env = (1.0 - u) * (1.0 - u);
        int c = 0;
        auto put = [&](double F, double gx, double gy, double gz) {
            // dF/dR = -dF/dd (d = point - R)
            F = keep ? F : 0.0;
            gx = keep ? -gx : 0.0;
            gy = keep ? -gy : 0.0;
            gz = keep ? -gz : 0.0;
            x[(int64_t)c * nb3 + p] = F;
            d[((int64_t)0 * nf + c) * nb3 + p] = gx;
            d[((int64_t)1 * nf + c) * nb3 + p] = gy;
            d[((int64_t)2 * nf + c) * nb3 + p] = gz;
            s2[c] += F * F;
            sd[c][0] += F * gx;
            sd[c][1] += F * gy;
            sd[c][2] += F * gz;
            ++c;
        };
        for (int e = 0; e < ns; ++e) {
            const double al = exps[e];
            const double ex = exp(-al * r2);
            const double q = ex * (-2.0 * al * env - 4.0 * (1.0 - u) / rc2);
            put(ex * env, q * ddx, q * ddy, q * ddz);
        }
        for (int e = 0; e < np; ++e) {
            const double al = exps[ns + e];
            const double ex = exp(-al * r2);
            const double q = ex * (-2.0 * al * env - 4.0 * (1.0 - u) / rc2);
            const double g = ex * env;
            put(ddx * g, ddx * q * ddx + g, ddx * q * ddy, ddx * q * ddz);
            put(ddy * g, ddy * q * ddx, ddy * q * ddy + g, ddy * q * ddz);
            put(ddz * g, ddz * q * ddx, ddz * q * ddy, ddz * q * ddz + g);
        }
    }
    // norms |F_c| = sqrt(sum F^2 dV) and their derivatives sum F dF dV / |F_c|
    for (int c = 0; c < nf; ++c) {
        const double a = block_sum_f(s2[c], red);
        const double b0 = block_sum_f(sd[c][0], red);
        const double b1 = block_sum_f(sd[c][1], red);
        const double b2 = block_sum_f(sd[c][2], red);
        if (threadIdx.x == 0) {
            const double n = sqrt(a * dV);
            nrm[c] = n;
            dn[c][0] = b0 * dV / n;
            dn[c][1] = b1 * dV / n;
            dn[c][2] = b2 * dV / n;
        }
    }
    __syncthreads();
    // x = F / n, dx/dR = dF / n - F dn / n^2
    for (int p = threadIdx.x; p < nb3; p += kFT)
        for (int c = 0; c < nf; ++c) {
            const double F = x[(int64_t)c * nb3 + p], n = nrm[c];
            x[(int64_t)c * nb3 + p] = F / n;
            for (int a = 0; a < 3; ++a) {
                double* q = d + ((int64_t)a * nf + c) * nb3 + p;
                *q = *q / n - F * dn[c][a] / (n * n);
            }
        }
    __syncthreads();
    // T x: 3 / dx^2 x - 1 / (2 dx^2) (sum of the six neighbours inside the box)
    const double cd = 3.0 / (dx * dx), co = -0.5 / (dx * dx);
    for (int p = threadIdx.x; p < nb3; p += kFT) {
        const int i = p / (nbox * nbox), j = (p / nbox) % nbox, k = p % nbox;
        for (int c = 0; c < nf; ++c) {
            const double* xc = x + (int64_t)c * nb3;
            double s = 0.0;
            if (i > 0) s += xc[p - nbox * nbox];
            if (i + 1 < nbox) s += xc[p + nbox * nbox];
            if (j > 0) s += xc[p - nbox];
            if (j + 1 < nbox) s += xc[p + nbox];
            if (k > 0) s += xc[p - 1];
            if (k + 1 < nbox) s += xc[p + 1];
            t[(int64_t)c * nb3 + p] = cd * xc[p] + co * s;
        }
    }
}

// Global gradient blocks of every force pair: G[pair][a nf + b] = 2 sum over its window
// entries (in window order) of g[off + a n_W + b] (spin factor 2), for gH and gS.
__global__ void k_force_gather(int npair, int nf, const int64_t* __restrict__ ent_ptr,
                               const int64_t* __restrict__ ent_off, const int32_t* __restrict__ ent_n,
                               const double* __restrict__ gH, const double* __restrict__ gS, double* __restrict__ GH,
                               double* __restrict__ GS) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nn = nf * nf;
    if (t >= (int64_t)npair * nn) return;
    const int64_t p = t / nn;
    const int e = (int)(t % nn), a = e / nf, b = e % nf;
    double h = 0.0, s = 0.0;
    for (int64_t q = ent_ptr[p]; q < ent_ptr[p + 1]; ++q) {
        const int64_t o = ent_off[q] + (int64_t)a * ent_n[q] + b;
        h += gH[o];
        s += gS[o];
    }
    GH[t] = 2.0 * h;
    GS[t] = 2.0 * s;
}

// Pair part for the force pairs (A <= B): over the points of A's box that lie in B's box,
//   f_A += sum_c dx_{A,c} sum_b [GH_cb ((T + v) x_{B,b}) + GS_cb x_{B,b}]
//   f_B += sum_b dx_{B,b} sum_c [GH_cb ((T + v) x_{A,c}) + GS_cb x_{A,c}]   (A != B)
// out[pair][0..2] = f_A, out[pair][3..5] = f_B.  One block per pair.
__global__ void __launch_bounds__(kFT) k_force_pairs(int npair, int nf, const int32_t* __restrict__ pA,
                                                      const int32_t* __restrict__ pB,
                                                      const int64_t* __restrict__ origin, int nbox, int n0, int n1,
                                                      int n2, const double* __restrict__ X,
                                                      const double* __restrict__ DX, const double* __restrict__ TX,
                                                      const double* __restrict__ v, const double* __restrict__ GH,
                                                      const double* __restrict__ GS, double* __restrict__ out) {
    __shared__ double red[kFT / 64];
    __shared__ double gh[64], gs[64];
    const int p = blockIdx.x;
    const int A = pA[p], B = pB[p];
    const int nn = nf * nf;
    for (int e = threadIdx.x; e < nn; e += kFT) {
        gh[e] = GH[(int64_t)p * nn + e];
        gs[e] = GS[(int64_t)p * nn + e];
    }
    __syncthreads();
    const int nb3 = nbox * nbox * nbox;
    const double* xa = X + (int64_t)A * nf * nb3;
    const double* xb = X + (int64_t)B * nf * nb3;
    const double* ta = TX + (int64_t)A * nf * nb3;
    const double* tb = TX + (int64_t)B * nf * nb3;
    const double* da = DX + (int64_t)A * 3 * nf * nb3;
    const double* db = DX + (int64_t)B * 3 * nf * nb3;
    const int64_t ax = origin[3 * A], ay = origin[3 * A + 1], az = origin[3 * A + 2];
    const int64_t bx = origin[3 * B], by = origin[3 * B + 1], bz = origin[3 * B + 2];
    double fa[3] = {0, 0, 0}, fb[3] = {0, 0, 0};
    for (int pa = threadIdx.x; pa < nb3; pa += kFT) {
        const int i = pa / (nbox * nbox), j = (pa / nbox) % nbox, k = pa % nbox;
        const int gx = wrap(ax + i, n0), gy = wrap(ay + j, n1), gz = wrap(az + k, n2);
        const int qi = wrap(gx - bx, n0), qj = wrap(gy - by, n1), qk = wrap(gz - bz, n2);
        if (qi >= nbox || qj >= nbox || qk >= nbox) continue;
        const int pb = (qi * nbox + qj) * nbox + qk;
        const double vg = v[((int64_t)gx * n1 + gy) * n2 + gz];
        double xA[8], yA[8], xB[8], yB[8];
        for (int c = 0; c < nf; ++c) {
            xA[c] = xa[(int64_t)c * nb3 + pa];
            yA[c] = ta[(int64_t)c * nb3 + pa] + vg * xA[c];
            xB[c] = xb[(int64_t)c * nb3 + pb];
            yB[c] = tb[(int64_t)c * nb3 + pb] + vg * xB[c];
        }
        for (int c = 0; c < nf; ++c) {
            double w = 0.0;
            for (int b = 0; b < nf; ++b) w += gh[c * nf + b] * yB[b] + gs[c * nf + b] * xB[b];
            for (int a = 0; a < 3; ++a) fa[a] += da[((int64_t)a * nf + c) * nb3 + pa] * w;
        }
        if (A != B) {
            for (int b = 0; b < nf; ++b) {
                double w = 0.0;
                for (int c = 0; c < nf; ++c) w += gh[c * nf + b] * yA[c] + gs[c * nf + b] * xA[c];
                for (int a = 0; a < 3; ++a) fb[a] += db[((int64_t)a * nf + b) * nb3 + pb] * w;
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        const double s = block_sum_f(fa[a], red);
        const double t = block_sum_f(fb[a], red);
        if (threadIdx.x == 0) {
            out[6 * (int64_t)p + a] = s;
            out[6 * (int64_t)p + 3 + a] = t;
        }
    }
}

// Q part: over the points g of A's cut-off sphere, owned by window B = owner[g] (row
// row_of[g] of the owned-row CSR, T-local columns), with A's first function at T-local
// position tpos[B natom + A]:  f_A += sum_c dx_{A,c}(g) sum_u GQ_B[c, u] x_u(g),
// GQ_B = 2 (F_B + beta D_B) (full k x k blocks, T ascending).  One block per atom.
__global__ void __launch_bounds__(kFT) k_force_q(int natom, int nf, const int64_t* __restrict__ origin, int nbox,
                                                  int n0, int n1, int n2, const double* __restrict__ X,
                                                  const double* __restrict__ DX, const int32_t* __restrict__ owner,
                                                  const int32_t* __restrict__ row_of, const int32_t* __restrict__ tpos,
                                                  const WinDesc* __restrict__ wins, const int64_t* __restrict__ xptr,
                                                  const int32_t* __restrict__ xidx, const double* __restrict__ xval,
                                                  const double* __restrict__ D, const double* __restrict__ F,
                                                  double beta, double* __restrict__ out) {
    __shared__ double red[kFT / 64];
    const int A = blockIdx.x;
    const int nb3 = nbox * nbox * nbox;
    const double* xa = X + (int64_t)A * nf * nb3;
    const double* da = DX + (int64_t)A * 3 * nf * nb3;
    const int64_t ax = origin[3 * A], ay = origin[3 * A + 1], az = origin[3 * A + 2];
    double f[3] = {0, 0, 0};
    for (int pa = threadIdx.x; pa < nb3; pa += kFT) {
        if (xa[pa] == 0.0) continue;              // outside the cut-off sphere (the first s function is > 0 inside)
        const int i = pa / (nbox * nbox), j = (pa / nbox) % nbox, k = pa % nbox;
        const int64_t g = ((int64_t)wrap(ax + i, n0) * n1 + wrap(ay + j, n1)) * n2 + wrap(az + k, n2);
        const int w = owner[g];
        const int ca = tpos[(int64_t)w * natom + A];
        if (ca < 0) continue;
        const WinDesc d = wins[w];
        const int64_t r = row_of[g];
        const double* Dw = D + d.kk_off;
        const double* Fw = F + d.kk_off;
        for (int c = 0; c < nf; ++c) {
            const int64_t rc = (int64_t)(ca + c) * d.k;
            double y = 0.0;
            for (int64_t e = xptr[r]; e < xptr[r + 1]; ++e) {
                const int u = xidx[e];
                y += 2.0 * (Fw[rc + u] + beta * Dw[rc + u]) * xval[e];
            }
            for (int a = 0; a < 3; ++a) f[a] += da[((int64_t)a * nf + c) * nb3 + pa] * y;
        }
    }
    for (int a = 0; a < 3; ++a) {
        const double s = block_sum_f(f[a], red);
        if (threadIdx.x == 0) out[3 * (int64_t)A + a] = s;
    }
}

// F_A = -2 dV (sum of A's pair parts, in pair order, + its Q part)
__global__ void k_force_total(int natom, const int64_t* __restrict__ aptr, const int64_t* __restrict__ aent,
                              const double* __restrict__ fpair, const double* __restrict__ fq, double dV,
                              double* __restrict__ Fout) {
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    if (A >= natom) return;
    double s[3] = {0, 0, 0};
    for (int64_t q = aptr[A]; q < aptr[A + 1]; ++q) {
        const int64_t e = aent[q];   // 6 pair + 3 side
        for (int a = 0; a < 3; ++a) s[a] += fpair[e + a];
    }
    for (int a = 0; a < 3; ++a) Fout[3 * A + a] = -2.0 * dV * (s[a] + fq[3 * A + a]);
}

}  // namespace
}  // namespace lres

using namespace lres;

hipError_t launch_atom_boxes(int natom, const double* pos, const int64_t* origin, int nbox, double dx,
                             const double* L, double rcut, const double* exps, int ns, int np, double dV, double* X,
                             double* DX, double* TX, hipStream_t st) {
    if (ns + 3 * np > 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_atom_boxes, dim3(natom), dim3(kFT), 0, st, natom, pos, origin, nbox, dx, L[0], L[1], L[2],
                       rcut, exps, ns, np, dV, X, DX, TX);
    return hipGetLastError();
}

hipError_t launch_force_gather(int npair, int nf, const int64_t* ent_ptr, const int64_t* ent_off,
                               const int32_t* ent_n, const double* gH, const double* gS, double* GH, double* GS,
                               hipStream_t st) {
    const int64_t n = (int64_t)npair * nf * nf;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_force_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, npair, nf, ent_ptr, ent_off,
                       ent_n, gH, gS, GH, GS);
    return hipGetLastError();
}

hipError_t launch_force_pairs(int npair, int nf, const int32_t* pA, const int32_t* pB, const int64_t* origin, int nbox,
                              const int64_t* shape, const double* X, const double* DX, const double* TX,
                              const double* v, const double* GH, const double* GS, double* out, hipStream_t st) {
    if (nf > 8) return hipErrorInvalidValue;
    if (npair == 0) return hipSuccess;
    hipLaunchKernelGGL(k_force_pairs, dim3(npair), dim3(kFT), 0, st, npair, nf, pA, pB, origin, nbox, (int)shape[0],
                       (int)shape[1], (int)shape[2], X, DX, TX, v, GH, GS, out);
    return hipGetLastError();
}

hipError_t launch_force_q(int natom, int nf, const int64_t* origin, int nbox, const int64_t* shape, const double* X,
                          const double* DX, const int32_t* owner, const int32_t* row_of, const int32_t* tpos,
                          const WinDesc* wins, const int64_t* xptr, const int32_t* xidx, const double* xval,
                          const double* D, const double* F, double beta, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_force_q, dim3(natom), dim3(kFT), 0, st, natom, nf, origin, nbox, (int)shape[0],
                       (int)shape[1], (int)shape[2], X, DX, owner, row_of, tpos, wins, xptr, xidx, xval, D, F, beta,
                       out);
    return hipGetLastError();
}

hipError_t launch_force_total(int natom, const int64_t* aptr, const int64_t* aent, const double* fpair,
                              const double* fq, double dV, double* F, hipStream_t st) {
    hipLaunchKernelGGL(k_force_total, dim3((natom + 255) / 256), dim3(256), 0, st, natom, aptr, aent, fpair, fq, dV,
                       F);
    return hipGetLastError();
}
