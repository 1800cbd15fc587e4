// The chemical-potential searches in reverse-communication form, shared by the host
// loops (lres_mu_solve, lres_mu_solve_warm) and the device controller of the queued
// warm density step (k_mu_begin, k_mu_step): the caller evaluates the electron count at the
// requested mu (one window pass) and feeds it back; every decision is taken here, so
// the host loop and the device-side search take the same steps in the same double
// arithmetic (no contraction of products into fma on either side).
//
// Warm search (lcao3d.py:412-448 with the pole route's passes): the count at the kept
// mu first (accepted within keep_tol, or lag), then secant steps from it with the last
// known slope dN/dmu and 20% overshoot until the root is bracketed, then Brent
// (R. P. Brent, Algorithms for Minimization without Derivatives, 1973, ch. 4; the
// reference's brentq, xtol 1e-12).  Where the bracket's high end lies far up the
// count's exponential rise towards the conduction band (f(hi) > 10 |f(lo)|) Brent runs
// on g = log(1 + f / A), A = 1.25 |f(lo)|, which has f's root and monotonicity there
// and is close to linear.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define LRES_HD __host__ __device__ inline

namespace lres {

enum RcStatus { kRcEval = 0, kRcDone = 1, kRcFail = 2 };

// Brent's root finder: inverse quadratic interpolation / secant steps safeguarded by
// bisection, tolerance 2 eps |b| + xtol / 2 as in the textbook; an evaluation with
// |f| <= ftol also ends it.  At most maxit evaluations.
struct BrentRC {
    double a, b, c, fa, fb, fc, d, e, xtol, ftol;
    int it, maxit;
};

// the loop body up to the next evaluation: kRcDone with the root in *out, or kRcEval
// with the next point in *out
LRES_HD int brent_top(BrentRC& s, double* out) {
#pragma clang fp contract(off)
    const double eps = 2.220446049250313e-16;
    if ((s.fb > 0 && s.fc > 0) || (s.fb < 0 && s.fc < 0)) {
        s.c = s.a;
        s.fc = s.fa;
        s.d = s.e = s.b - s.a;
    }
    if (fabs(s.fc) < fabs(s.fb)) {
        s.a = s.b;
        s.b = s.c;
        s.c = s.a;
        s.fa = s.fb;
        s.fb = s.fc;
        s.fc = s.fa;
    }
    const double tol = 2.0 * eps * fabs(s.b) + 0.5 * s.xtol;
    const double mid = 0.5 * (s.c - s.b);
    if (fabs(mid) <= tol || fabs(s.fb) <= s.ftol) {
        *out = s.b;
        return kRcDone;
    }
    if (fabs(s.e) >= tol && fabs(s.fa) > fabs(s.fb)) {
        double p, q, r;
        const double sr = s.fb / s.fa;
        if (s.a == s.c) {
            p = 2.0 * mid * sr;
            q = 1.0 - sr;
        } else {
            q = s.fa / s.fc;
            r = s.fb / s.fc;
            p = sr * (2.0 * mid * q * (q - r) - (s.b - s.a) * (r - 1.0));
            q = (q - 1.0) * (r - 1.0) * (sr - 1.0);
        }
        if (p > 0)
            q = -q;
        else
            p = -p;
        const double lim1 = 3.0 * mid * q - fabs(tol * q), lim2 = fabs(s.e * q);
        if (2.0 * p < (lim1 < lim2 ? lim1 : lim2)) {
            s.e = s.d;
            s.d = p / q;
        } else {
            s.d = mid;
            s.e = s.d;
        }
    } else {
        s.d = mid;
        s.e = s.d;
    }
    s.a = s.b;
    s.fa = s.fb;
    s.b += (fabs(s.d) > tol) ? s.d : (mid > 0 ? tol : -tol);
    *out = s.b;
    return kRcEval;
}

// start on [a, b] with f(a) = fa, f(b) = fb: kRcFail if they have the same sign
LRES_HD int brent_begin(BrentRC& s, double a, double b, double fa, double fb, double xtol, double ftol, int maxit,
                        double* out) {
    if (fa * fb > 0.0) return kRcFail;
    s.a = a;
    s.b = b;
    s.fa = fa;
    s.fb = fb;
    s.xtol = xtol;
    s.ftol = ftol;
    s.it = 0;
    s.maxit = maxit;
    if (fabs(fa) <= ftol) {
        *out = a;
        return kRcDone;
    }
    if (fabs(fb) <= ftol) {
        *out = b;
        return kRcDone;
    }
    s.c = a;
    s.fc = fa;
    s.d = b - a;
    s.e = s.d;
    return brent_top(s, out);
}

// f at the point last returned by brent_top
LRES_HD int brent_feed(BrentRC& s, double f, double* out) {
    s.fb = f;
    if (++s.it >= s.maxit) return kRcFail;
    return brent_top(s, out);
}

// Warm search states
enum WarmPhase {
    kWarmKeep = 0,      // the pass in flight is the keep test at mu0
    kWarmSecant = 1,    // a secant trial
    kWarmBrent = 2,     // a Brent trial
    kWarmDone = 3,      // mu accepted
    kWarmNoBracket = 4, // no bracket within 60 secant steps (or the Brent bracket broke)
    kWarmNoConv = 5,    // Brent did not converge in 200 evaluations
    kWarmEdge = 6       // the keep-edge inertia test failed: a new filter edge is needed
};
// how a finished search ended
enum WarmOutcome { kOutKeep = 0, kOutLag = 1, kOutResolve = 2 };

struct WarmSearch {
    // inputs
    double N, e_lo, ftol, xtol, keep_tol;
    int lag;
    // state
    int phase, outcome, evals, sec_it, logmode, by_brent;
    double mu0, slope, a, fa, step, x, Ab;
    double last_mu, last_f, last_nsum, last_esum, mu;
    BrentRC br;
};

LRES_HD bool warm_terminal(const WarmSearch& s) { return s.phase >= kWarmDone; }

// propose the next secant trial from (a, fa) and step
LRES_HD void warm_secant_next(WarmSearch& s) {
#pragma clang fp contract(off)
    if (s.sec_it >= 60) {
        s.phase = kWarmNoBracket;
        return;
    }
    ++s.sec_it;
    s.step = fmin(fmax(s.step, -0.25), 0.25);
    if (fabs(s.step) < 1e-12) s.step = s.fa < 0 ? 1e-12 : -1e-12;
    double b = s.a + s.step;
    if (b <= s.e_lo + 1e-6) b = 0.5 * (s.a + s.e_lo + 1e-6);
    s.x = b;
    s.phase = kWarmSecant;
}

LRES_HD void warm_accept(WarmSearch& s, double mu, int outcome) {
    s.mu = mu;
    s.outcome = outcome;
    s.phase = kWarmDone;
}

// secant phase from (mu0, f0) (the count there known)
LRES_HD void warm_start_secant(WarmSearch& s, double mu0, double f0) {
#pragma clang fp contract(off)
    s.a = mu0;
    s.fa = f0;
    s.step = s.slope != 0.0 ? -f0 / s.slope : (f0 < 0 ? 1e-3 : -1e-3);
    s.sec_it = 0;
    warm_secant_next(s);
}

LRES_HD void warm_init(WarmSearch& s, double N, double e_lo, double mu0, double slope, double keep_tol,
                       double ftol, double xtol, int lag) {
    s.N = N;
    s.e_lo = e_lo;
    s.ftol = ftol;
    s.xtol = xtol;
    s.keep_tol = keep_tol;
    s.lag = lag;
    s.outcome = kOutResolve;
    s.evals = 0;
    s.sec_it = 0;
    s.logmode = 0;
    s.by_brent = 0;
    s.mu0 = mu0;
    s.slope = slope;
    s.a = s.fa = s.step = s.Ab = 0.0;
    s.last_mu = s.last_f = s.last_nsum = s.last_esum = s.mu = NAN;
}

// keep test first: the first trial is mu0 itself
LRES_HD void warm_begin_keep(WarmSearch& s, double N, double e_lo, double mu0, double slope, double keep_tol,
                             double ftol, double xtol, int lag) {
    warm_init(s, N, e_lo, mu0, slope, keep_tol, ftol, xtol, lag);
    s.x = mu0;
    s.phase = kWarmKeep;
}

// the count f0 at mu0 already known (and above the tolerance): secant steps at once
LRES_HD void warm_begin(WarmSearch& s, double N, double e_lo, double mu0, double f0, double slope, double ftol,
                        double xtol) {
    warm_init(s, N, e_lo, mu0, slope, 0.0, ftol, xtol, 0);
    warm_start_secant(s, mu0, f0);
}

// the pass at s.x gave (sum N_B, sum E_B)
LRES_HD void warm_feed(WarmSearch& s, double nsum, double esum) {
#pragma clang fp contract(off)
    const double f = 2.0 * nsum - s.N;
    ++s.evals;
    s.last_mu = s.x;
    s.last_f = f;
    s.last_nsum = nsum;
    s.last_esum = esum;
    if (s.phase == kWarmKeep) {
        if (fabs(f) <= s.keep_tol) {
            warm_accept(s, s.x, kOutKeep);
        } else if (s.lag && fabs(f) < 0.25) {
            warm_accept(s, s.x, kOutLag);
        } else {
            warm_start_secant(s, s.x, f);
        }
        return;
    }
    if (s.phase == kWarmSecant) {
        const double b = s.x, fb = f;
        // slope of the count between the two points when positive and finite
        if (b != s.a && isfinite(fb) && (fb - s.fa) / (b - s.a) > 0) s.slope = (fb - s.fa) / (b - s.a);
        if (fabs(fb) <= s.ftol) {
            warm_accept(s, b, kOutResolve);
            return;
        }
        if ((s.fa < 0) != (fb < 0)) {
            const bool lo_a = s.a < b;
            const double lo = lo_a ? s.a : b, hi = lo_a ? b : s.a, flo = lo_a ? s.fa : fb, fhi = lo_a ? fb : s.fa;
            double out = 0.0;
            int r;
            s.by_brent = 1;
            if (fhi > 10.0 * fabs(flo)) {
                s.logmode = 1;
                s.Ab = 1.25 * fabs(flo);
                r = brent_begin(s.br, lo, hi, log1p(flo / s.Ab), log1p(fhi / s.Ab), s.xtol, log1p(s.ftol / s.Ab), 200,
                                &out);
            } else {
                s.logmode = 0;
                r = brent_begin(s.br, lo, hi, flo, fhi, s.xtol, s.ftol, 200, &out);
            }
            if (r == kRcDone) {
                warm_accept(s, out, kOutResolve);
            } else if (r == kRcEval) {
                s.x = out;
                s.phase = kWarmBrent;
            } else {
                s.phase = kWarmNoBracket;
            }
            return;
        }
        const double sl = (fb - s.fa) / (b - s.a);
        // same side of the root: secant extrapolation with 20% overshoot to bracket it
        s.step = sl > 0 ? -1.2 * fb / sl : 4.0 * s.step;
        s.a = b;
        s.fa = fb;
        warm_secant_next(s);
        return;
    }
    if (s.phase == kWarmBrent) {
        double out = 0.0;
        const int r = brent_feed(s.br, s.logmode ? log1p(f / s.Ab) : f, &out);
        if (r == kRcDone) {
            warm_accept(s, out, kOutResolve);
        } else if (r == kRcEval) {
            s.x = out;
        } else {
            s.phase = kWarmNoConv;
        }
    }
}

// Device bookkeeping of the queued warm density step: the search plus whether the
// accepted mu needs a pass of its own (a Brent root that was not the last trial).
struct StepState {
    WarmSearch s;
    int final_pass;   // the final pass is queued at the accepted mu
    int final_ran;    // final passes that ran
};

// stages of the controller between the queued passes of a device step: a trial slot's
// sums fed to the search (kStepFeed; kStepFeedFinal also decides the final slot: a pass
// at the accepted mu when it was not the last trial, and the density either way), the
// final slot's sums recorded and the state published to the host (kStepPublish), the
// next batch's first trial (kStepResume)
enum StepStage { kStepFeed = 0, kStepFeedFinal = 1, kStepPublish = 2, kStepResume = 3 };

}  // namespace lres
