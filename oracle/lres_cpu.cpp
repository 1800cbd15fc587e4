// CPU restatement of the LRes window pass, for timing the CPU baseline and as a
// second checker.  TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's
// cpu_baseline leg through oracle/cpu_native.py, never by the product package.
//
// lres_cpu_eig_pass  — the reference's eigen route, lrft_windows
//                      (src/lrft_cpp.cpp:129-233): per window a dense
//                      generalised eigendecomposition of (H_W, S_W) (dsygvd for
//                      the full spectrum, dsygvx for the part below `upper`,
//                      src/lrft_cpp.cpp:79-114), Y = X[own, W] Z, c_k = sum Y^2 dV,
//                      density sum_k r_M(lam_k) Y^2 on the owned points and the
//                      float32 Y^2 store (src/lrft_cpp.cpp:164-192).
// lres_cpu_pole_pass — the pole route of benchmarks/window_resolvent_vs_eig.py:58-65:
//                      per window and upper-half-plane pole one complex LU solve
//                      (zgesv) of (z_j S_W - H_W) G = E_T, E_T the unit columns of
//                      the functions touching the owned points, P = 2 Re sum w_j G,
//                      F = 2 Re sum w_j z_j G, then n(r) = x_r^T P x_r, N_B, E_B.
// OpenMP over windows (dynamic schedule); LAPACK comes as function pointers
// from the caller (scipy's cython_lapack), so each call is single-threaded
// LAPACK on one OpenMP thread, as src/lrft_cpp.cpp:143-152 arranges.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

typedef std::complex<double> zc;
typedef void (*dsygvd_fn)(int*, char*, char*, int*, double*, int*, double*, int*, double*, double*, int*, int*,
                          int*, int*);
typedef void (*dsygvx_fn)(int*, char*, char*, char*, int*, double*, int*, double*, int*, double*, double*, int*,
                          int*, double*, int*, double*, double*, int*, double*, int*, int*, int*, int*);
typedef void (*zgesv_fn)(int*, int*, zc*, int*, int*, zc*, int*, int*);

namespace {

// 1 / (1 + ((lam - c) / R)^M), 0 where the power overflows (three_d.py:221-225)
double occ(double lam, double c, double R, int M) {
    const double p = std::pow(std::fabs((lam - c) / R), M);
    return std::isfinite(p) ? 1.0 / (1.0 + p) : 0.0;
}

// A[W,W] of a CSR matrix, column-major n x n, through the global -> local map pos
void gather(const int64_t* p, const int64_t* ix, const double* v, const int64_t* W, int n, const int32_t* pos,
            double* out) {
    std::fill(out, out + (size_t)n * n, 0.0);
    for (int a = 0; a < n; ++a) {
        const int64_t row = W[a];
        for (int64_t e = p[row]; e < p[row + 1]; ++e) {
            const int b = pos[ix[e]];
            if (b >= 0) out[(size_t)b * n + a] = v[e];
        }
    }
}

struct Csr {
    const int64_t* p;
    const int64_t* i;
    const double* v;
};

}  // namespace

extern "C" {

// Eigen-route pass over all windows.  Outputs (caller-allocated):
//   lam, c: sum_w n_W doubles each, window w from Woff[w] - Woff[0]; m[w] eigenpairs
//   kept (n_W with upper = inf); dens (ng, may be NULL) on owned points;
//   y2 (may be NULL): float32 Y^2 of window w at y2off[w], nown_w x n_W row-major
//   (only the first m[w] columns of each row are meaningful).
// Returns 0, or the LAPACK info (negative: argument error, positive: failure).
int lres_cpu_eig_pass(void* dsygvd_ptr, void* dsygvx_ptr, const int64_t* Hp, const int64_t* Hi, const double* Hv,
                      const int64_t* Sp, const int64_t* Si, const double* Sv, const int64_t* Xp, const int64_t* Xi,
                      const double* Xv, const int64_t* Wflat, const int64_t* Woff, int64_t nwin,
                      const int64_t* Oflat, const int64_t* Ooff, int64_t nb, double dV, double mu, double e_lo, int M,
                      double upper, int want_density, int nthreads, double* lam, double* c, int32_t* m,
                      double* dens, float* y2, const int64_t* y2off) {
    dsygvd_fn dsygvd = (dsygvd_fn)dsygvd_ptr;
    dsygvx_fn dsygvx = (dsygvx_fn)dsygvx_ptr;
    const Csr H{Hp, Hi, Hv}, S{Sp, Si, Sv}, X{Xp, Xi, Xv};
    const double cc = 0.5 * (mu + e_lo), R = 0.5 * (mu - e_lo);
    const bool full = !std::isfinite(upper);
    int err = 0;
#pragma omp parallel num_threads(std::max(1, nthreads))
    {
        std::vector<int32_t> pos(nb, -1);
        std::vector<double> A, B, Z, w, work, Y;
        std::vector<int> iwork, ifail;
#pragma omp for schedule(dynamic)
        for (int64_t b = 0; b < nwin; ++b) {
            const int64_t* W = Wflat + Woff[b];
            const int n = (int)(Woff[b + 1] - Woff[b]);
            const int64_t l0 = Woff[b] - Woff[0];
            if (n == 0) {
                m[b] = 0;
                continue;
            }
            for (int a = 0; a < n; ++a) pos[W[a]] = a;
            A.resize((size_t)n * n);
            B.resize((size_t)n * n);
            Z.resize((size_t)n * n);
            w.resize(n);
            gather(H.p, H.i, H.v, W, n, pos.data(), A.data());
            gather(S.p, S.i, S.v, W, n, pos.data(), B.data());
            int itype = 1, info = 0, nn = n, lwork = -1, liwork = -1, mm = 0;
            char jobz = 'V', uplo = 'U', range = 'V';
            double wq = 0.0;
            int iq = 0;
            if (full) {
                dsygvd(&itype, &jobz, &uplo, &nn, A.data(), &nn, B.data(), &nn, w.data(), &wq, &lwork, &iq, &liwork,
                       &info);
                lwork = (int)wq;
                liwork = iq;
                work.resize(std::max(1, lwork));
                iwork.resize(std::max(1, liwork));
                dsygvd(&itype, &jobz, &uplo, &nn, A.data(), &nn, B.data(), &nn, w.data(), work.data(), &lwork,
                       iwork.data(), &liwork, &info);
                mm = n;
                std::memcpy(Z.data(), A.data(), sizeof(double) * (size_t)n * n);
            } else {
                double vl = -1e300, vu = upper, abstol = 0.0;
                int il = 0, iu = 0;
                ifail.resize(n);
                iwork.resize(5 * (size_t)n);
                dsygvx(&itype, &jobz, &range, &uplo, &nn, A.data(), &nn, B.data(), &nn, &vl, &vu, &il, &iu, &abstol,
                       &mm, w.data(), Z.data(), &nn, &wq, &lwork, iwork.data(), ifail.data(), &info);
                lwork = (int)wq;
                work.resize(std::max(1, lwork));
                dsygvx(&itype, &jobz, &range, &uplo, &nn, A.data(), &nn, B.data(), &nn, &vl, &vu, &il, &iu, &abstol,
                       &mm, w.data(), Z.data(), &nn, work.data(), &lwork, iwork.data(), ifail.data(), &info);
            }
            if (info != 0 && (full || info < 0)) {
#pragma omp critical
                err = info;
                for (int a = 0; a < n; ++a) pos[W[a]] = -1;
                continue;
            }
            m[b] = mm;
            // Y = X[own, W] Z (Z column-major: eigenvector k is column k)
            const int64_t o0 = Ooff[b], no = Ooff[b + 1] - Ooff[b];
            Y.assign((size_t)no * std::max(mm, 1), 0.0);
            for (int64_t r = 0; r < no; ++r) {
                const int64_t g = Oflat[o0 + r];
                double* y = Y.data() + (size_t)r * mm;
                for (int64_t e = X.p[g]; e < X.p[g + 1]; ++e) {
                    const int a = pos[X.i[e]];
                    if (a < 0) continue;
                    const double x = X.v[e];
                    for (int k = 0; k < mm; ++k) y[k] += x * Z[(size_t)k * n + a];
                }
            }
            for (int k = 0; k < mm; ++k) {
                double s = 0.0;
                for (int64_t r = 0; r < no; ++r) s += Y[(size_t)r * mm + k] * Y[(size_t)r * mm + k];
                lam[l0 + k] = w[k];
                c[l0 + k] = s * dV;
            }
            if (want_density && dens) {
                std::vector<double> oc(mm);
                for (int k = 0; k < mm; ++k) oc[k] = occ(w[k], cc, R, M);
                for (int64_t r = 0; r < no; ++r) {
                    double s = 0.0;
                    for (int k = 0; k < mm; ++k) {
                        const double y = Y[(size_t)r * mm + k];
                        s += oc[k] * y * y;
                    }
                    dens[Oflat[o0 + r]] = s;
                }
            }
            if (y2) {
                float* out = y2 + y2off[b];
                for (int64_t r = 0; r < no; ++r)
                    for (int k = 0; k < mm; ++k) {
                        const double y = Y[(size_t)r * mm + k];
                        out[(size_t)r * n + k] = (float)(y * y);
                    }
            }
            for (int a = 0; a < n; ++a) pos[W[a]] = -1;
        }
    }
    return err;
}

// Pole-route pass over all windows: dens (ng, may be NULL), NB, EB (nwin), no spin factor.
int lres_cpu_pole_pass(void* zgesv_ptr, const int64_t* Hp, const int64_t* Hi, const double* Hv, const int64_t* Sp,
                       const int64_t* Si, const double* Sv, const int64_t* Xp, const int64_t* Xi, const double* Xv,
                       const int64_t* Wflat, const int64_t* Woff, int64_t nwin, const int64_t* Oflat,
                       const int64_t* Ooff, int64_t nb, double dV, double mu, double e_lo, int M, int nthreads,
                       double* dens, double* NB, double* EB) {
    zgesv_fn zgesv = (zgesv_fn)zgesv_ptr;
    const Csr H{Hp, Hi, Hv}, S{Sp, Si, Sv}, X{Xp, Xi, Xv};
    const double cc = 0.5 * (mu + e_lo), R = 0.5 * (mu - e_lo);
    const double kPi = 3.14159265358979323846;
    int err = 0;
#pragma omp parallel num_threads(std::max(1, nthreads))
    {
        std::vector<int32_t> pos(nb, -1);
        std::vector<double> Hd, Sd, P, F;
        std::vector<zc> A, G;
        std::vector<int> ipiv, tloc;
        std::vector<char> touch;
#pragma omp for schedule(dynamic)
        for (int64_t b = 0; b < nwin; ++b) {
            const int64_t* W = Wflat + Woff[b];
            const int n = (int)(Woff[b + 1] - Woff[b]);
            const int64_t o0 = Ooff[b], no = Ooff[b + 1] - Ooff[b];
            NB[b] = EB[b] = 0.0;
            if (n == 0) continue;
            for (int a = 0; a < n; ++a) pos[W[a]] = a;
            // T: window functions with support on the owned points, ascending
            touch.assign(n, 0);
            for (int64_t r = 0; r < no; ++r) {
                const int64_t g = Oflat[o0 + r];
                for (int64_t e = X.p[g]; e < X.p[g + 1]; ++e)
                    if (pos[X.i[e]] >= 0) touch[pos[X.i[e]]] = 1;
            }
            std::vector<int> T;
            tloc.assign(n, -1);
            for (int a = 0; a < n; ++a)
                if (touch[a]) {
                    tloc[a] = (int)T.size();
                    T.push_back(a);
                }
            const int k = (int)T.size();
            Hd.resize((size_t)n * n);
            Sd.resize((size_t)n * n);
            gather(H.p, H.i, H.v, W, n, pos.data(), Hd.data());
            gather(S.p, S.i, S.v, W, n, pos.data(), Sd.data());
            P.assign((size_t)k * k, 0.0);
            F.assign((size_t)k * k, 0.0);
            A.resize((size_t)n * n);
            G.resize((size_t)n * k);
            ipiv.resize(n);
            for (int j = 0; j < M / 2; ++j) {
                const double th = kPi * (2 * j + 1) / M;
                const zc e(std::cos(th), std::sin(th));
                const zc z = cc + R * e, wj = R * e / (double)M, wz = wj * z;
                for (size_t q = 0; q < (size_t)n * n; ++q) A[q] = z * Sd[q] - Hd[q];
                std::fill(G.begin(), G.end(), zc(0.0, 0.0));
                for (int t = 0; t < k; ++t) G[(size_t)t * n + T[t]] = 1.0;
                int nn = n, nrhs = k, info = 0;
                zgesv(&nn, &nrhs, A.data(), &nn, ipiv.data(), G.data(), &nn, &info);
                if (info != 0) {
#pragma omp critical
                    err = info;
                    break;
                }
                // P[s][t] += 2 Re(w G[T_s, t]) (G column t solves for unit column T_t)
                for (int t = 0; t < k; ++t)
                    for (int s = 0; s < k; ++s) {
                        const zc g = G[(size_t)t * n + T[s]];
                        P[(size_t)s * k + t] += 2.0 * (wj * g).real();
                        F[(size_t)s * k + t] += 2.0 * (wz * g).real();
                    }
            }
            // n(r) = x_T^T P x_T on the owned points, N_B and E_B
            std::vector<double> xv(k);
            double sn = 0.0, se = 0.0;
            for (int64_t r = 0; r < no; ++r) {
                const int64_t g = Oflat[o0 + r];
                std::vector<std::pair<int, double>> ent;
                for (int64_t e = X.p[g]; e < X.p[g + 1]; ++e) {
                    const int a = pos[X.i[e]];
                    if (a >= 0 && tloc[a] >= 0) ent.push_back({tloc[a], X.v[e]});
                }
                double nr = 0.0, er = 0.0;
                for (auto& u : ent)
                    for (auto& v2 : ent) {
                        nr += u.second * P[(size_t)u.first * k + v2.first] * v2.second;
                        er += u.second * F[(size_t)u.first * k + v2.first] * v2.second;
                    }
                if (dens) dens[g] = nr;
                sn += nr;
                se += er;
            }
            NB[b] = sn * dV;
            EB[b] = se * dV;
            for (int a = 0; a < n; ++a) pos[W[a]] = -1;
        }
    }
    return err;
}

int lres_cpu_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
