// Host-only driver of the reverse-communication mu searches (lresdft_amd/csrc/lres_musearch.h)
// for the CPU tests: the count comes from a Python callback.
#include <hip/hip_runtime.h>

#include "../../lresdft_amd/csrc/lres_musearch.h"

typedef double (*count_fn)(double mu);

extern "C" {

// Warm search as lres_step_enqueue runs it (keep_first != 0: keep test at mu0 first)
// or as lres_mu_solve_warm runs it (keep_first == 0, f0 the known count at mu0).
// out[8] = {phase, outcome, evals, mu, slope, by_brent, logmode, last_mu}
int sim_warm(double N, double e_lo, double mu0, double f0, double slope, double keep_tol, double ftol, double xtol,
             int lag, int keep_first, count_fn f, double* out) {
    lres::WarmSearch s;
    if (keep_first)
        lres::warm_begin_keep(s, N, e_lo, mu0, slope, keep_tol, ftol, xtol, lag);
    else
        lres::warm_begin(s, N, e_lo, mu0, f0, slope, ftol, xtol);
    while (!lres::warm_terminal(s)) {
        const double c = f(s.x);               // count, f = 2 nsum - N
        lres::warm_feed(s, 0.5 * (c + N), 0.0);
    }
    out[0] = s.phase;
    out[1] = s.outcome;
    out[2] = s.evals;
    out[3] = s.mu;
    out[4] = s.slope;
    out[5] = s.by_brent;
    out[6] = s.logmode;
    out[7] = s.last_mu;
    return s.phase;
}

// Brent on [a, b]: returns the status, root in *root, evaluations in *evals
int sim_brent(double a, double b, double xtol, double ftol, count_fn f, double* root, int* evals) {
    lres::BrentRC br;
    double x = 0.0;
    *evals = 2;
    int r = lres::brent_begin(br, a, b, f(a), f(b), xtol, ftol, 200, &x);
    while (r == lres::kRcEval) {
        ++*evals;
        r = lres::brent_feed(br, f(x), &x);
    }
    *root = x;
    return r;
}
}
