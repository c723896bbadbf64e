/* CFAR threshold design (reference include/atkdsp.h:1733-1860, section 19 of
 * the detection track; the reference solver is src/cfar.c:23-268): the
 * defaults, the equivalent number of independent reference cells for a
 * window-correlated periodogram, that correlation for a window, and the
 * threshold factor alpha for a false-alarm rate. Host code: alpha is a
 * design-time constant, made once per configuration.
 *
 * The solver is this library's own. A cell under test X and the reference
 * statistic T are unit-mean K-look gamma variables (T built from N cells, or
 * the n_eff independent cells they are worth); a cell hits when X > alpha T,
 * so
 *     Pfa(alpha) = P(T < X / alpha) = E_X[ F_T(X / alpha) ].
 * F_T is known in closed form for every mode:
 *     CA  the mean of n_eff cells: gamma of shape n_eff K (and then Pfa itself
 *         is a regularized incomplete beta: I_{1/(1+alpha/n_eff)}(n_eff K, K));
 *     GO  the larger of two side means of shape n_eff K / 2: F_side^2;
 *     SO  the smaller: 1 - (1 - F_side)^2;
 *     OS  the k-th smallest of N cells: I_{F(y)}(k, N - k + 1), the order
 *         statistic's distribution through the regularized incomplete beta.
 * For GO, SO and OS the expectation over X is integrated in u = ln x with
 * 16-point Gauss-Legendre panels of width at most 1/32 over the span where
 * the integrand is within e^-60 of its peak, in log space throughout; alpha
 * is then the root of ln Pfa(alpha) = ln pfa in ln alpha (bisection to a
 * bracket, then Brent). Accurate to about 1e-13 relative where the
 * reference's Simpson grid is accurate to about 1e-9, so the two agree to the
 * reference's accuracy. With correlated cells, OS is solved at the rate the
 * independent-cell CA test has at the correlated CA threshold, as the
 * reference defines it (include/atkdsp.h:1802-1817). */
#include <math.h>
#include <stddef.h>
#include <string.h>
#include "internal.h"

#define CFAR_RHO_MAX 4

/* the layout the reference header compiles to (96 bytes; its comment says 104) */
_Static_assert(sizeof(atkdsp_cfar_config) == 96 && offsetof(atkdsp_cfar_config, pfa) == 16 &&
                   offsetof(atkdsp_cfar_config, rho2) == 64, "the reference's configuration layout");

void atkdsp_cfar_defaults(atkdsp_cfar_config *c) {
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->mode = ATKDSP_CFAR_CA;
    c->n_ref = 16;
    c->n_guard = 4;
    c->k_frames = 1;
    c->pfa = 1e-6;
    c->looks = 1.0;
    c->os_k = 0;
    c->open_m = 2;
    c->open_n = 3;
    c->close_k = 3;
    c->merge_gap = 1;
    c->min_bins = 1;
    c->max_tracks = 256;
}

double atkdsp_cfar_n_eff(int n_ref, const double *rho2) {
    if (n_ref < 1) return 0.0;
    const double N = 2.0 * (double)n_ref;
    if (!rho2) return N;
    /* pairs m bins apart inside each side's n_ref contiguous cells */
    double pairs = 0.0;
    for (int m = 1; m <= CFAR_RHO_MAX && m < n_ref; ++m) pairs += 2.0 * (double)(n_ref - m) * rho2[m - 1];
    return N / (1.0 + 2.0 * pairs / N);
}

int atkdsp_cfar_window_rho2(const float *window, size_t n, double *rho2) {
    if (!window || !rho2 || n < 2) return ATKDSP_E_ARG;
    double e = 0.0;
    for (size_t i = 0; i < n; ++i) e += (double)window[i] * (double)window[i];
    if (!(e > 0.0)) return ATKDSP_E_ARG;
    for (int m = 1; m <= CFAR_RHO_MAX; ++m) {
        double re = 0.0, im = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double w2 = (double)window[i] * (double)window[i];
            const double ph = 2.0 * ATK_PI * (double)m * (double)i / (double)n;
            re += w2 * cos(ph);
            im -= w2 * sin(ph);
        }
        rho2[m - 1] = (re * re + im * im) / (e * e);
    }
    return ATKDSP_OK;
}

/* ---- log-space special functions ------------------------------------------------ */

static double log1m_exp(double a) {                         /* ln(1 - e^a), a <= 0 */
    if (a == -INFINITY) return 0.0;
    return a > -0.693147180559945309 ? log(-expm1(a)) : log1p(-exp(a));
}

/* ln P(a, x) and ln Q(a, x), the regularized lower and upper incomplete
 * gamma: the power series below a + 1, Legendre's continued fraction (by
 * the modified Lentz method) above */
static void lgam_pq(double a, double x, double *lp, double *lq) {
    if (!(x > 0.0)) { *lp = -INFINITY; *lq = 0.0; return; }
    if (x == INFINITY) { *lp = 0.0; *lq = -INFINITY; return; }
    const double lead = a * log(x) - x - lgamma(a);
    if (x < a + 1.0) {
        double term = 1.0 / a, sum = term, ap = a;
        for (int n = 0; n < 1000000; ++n) {
            ap += 1.0;
            term *= x / ap;
            sum += term;
            if (term < sum * 1e-17) break;
        }
        *lp = lead + log(sum);
        *lq = log1m_exp(*lp);
        return;
    }
    const double tiny = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 1000000; ++i) {
        const double an = -(double)i * ((double)i - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < 1e-16) break;
    }
    *lq = lead + log(h);
    *lp = log1m_exp(*lq);
}

void atk_cfar_lgam_pq(double a, double x, double *lp, double *lq) { lgam_pq(a, x, lp, lq); }

/* the continued fraction of the incomplete beta (modified Lentz) */
static double beta_cf(double a, double b, double x) {
    const double tiny = 1e-300;
    const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0, d = 1.0 - qab * x / qap;
    if (fabs(d) < tiny) d = tiny;
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m < 1000000; ++m) {
        const double m2 = 2.0 * m;
        double aa = (double)m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < tiny) d = tiny;
        c = 1.0 + aa / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        h *= d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < tiny) d = tiny;
        c = 1.0 + aa / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < 1e-16) break;
    }
    return h;
}

/* ln I_x(a, b) and ln(1 - I_x(a, b)) from ln x and ln(1 - x) (so that x may
 * be far below the smallest double) */
static void lbeta_i(double a, double b, double lx, double l1x, double *li, double *l1i) {
    if (lx == -INFINITY) { *li = -INFINITY; *l1i = 0.0; return; }
    if (l1x == -INFINITY) { *li = 0.0; *l1i = -INFINITY; return; }
    const double x = exp(lx);
    const double lfront = lgamma(a + b) - lgamma(a) - lgamma(b) + a * lx + b * l1x;
    if (x < (a + 1.0) / (a + b + 2.0)) {
        *li = lfront + log(beta_cf(a, b, x)) - log(a);
        *l1i = log1m_exp(*li);
    } else {
        *l1i = lfront + log(beta_cf(b, a, exp(l1x))) - log(b);
        *li = log1m_exp(*l1i);
    }
}

/* ---- the statistic T and the false-alarm rate --------------------------------- */

typedef struct {
    int mode, N, k;      /* OS: the k-th smallest of N */
    double ne, K;        /* effective cells, looks per cell */
} tstat;

/* ln F_T(y), ln(1 - F_T(y)) */
static void log_ft(const tstat *t, double y, double *lf, double *lg) {
    double lp, lq;
    switch (t->mode) {
    case ATKDSP_CFAR_CA: {
        const double s = t->ne * t->K;
        lgam_pq(s, s * y, lf, lg);
        return;
    }
    case ATKDSP_CFAR_GO: {
        const double s = 0.5 * t->ne * t->K;
        lgam_pq(s, s * y, &lp, &lq);
        *lf = 2.0 * lp;
        *lg = log1m_exp(*lf);
        return;
    }
    case ATKDSP_CFAR_SO: {
        const double s = 0.5 * t->ne * t->K;
        lgam_pq(s, s * y, &lp, &lq);
        *lg = 2.0 * lq;
        *lf = log1m_exp(*lg);
        return;
    }
    default:
        lgam_pq(t->K, t->K * y, &lp, &lq);
        lbeta_i((double)t->k, (double)(t->N - t->k + 1), lp, lq, lf, lg);
    }
}

/* ln of the integrand of Pfa in u = ln x: ln f_X(x) + u + ln F_T(x / alpha) */
static double log_integrand(const tstat *t, double lalpha, double u) {
    const double K = t->K, x = exp(u);
    const double lfx = K * log(K) + (K - 1.0) * u - K * x - lgamma(K);
    double lf, lg;
    log_ft(t, exp(u - lalpha), &lf, &lg);
    return lfx + u + lf;
}

/* 16-point Gauss-Legendre on [-1, 1]: the positive nodes and their weights */
static const double GL16_X[8] = {0.09501250983763745, 0.2816035507792589, 0.45801677765722737, 0.6178762444026438, 0.755404408355003, 0.8656312023878318, 0.9445750230732326, 0.9894009349916499};
static const double GL16_W[8] = {0.18945061045506859, 0.1826034150449236, 0.16915651939500262, 0.14959598881657676, 0.12462897125553403, 0.09515851168249259, 0.062253523938647706, 0.027152459411754037};

/* ln of a positive integral in u of e^{lg(u)} over the span where lg is
 * within 60 of its maximum */
static double log_integral(const tstat *t, double lalpha) {
    const double ulo = -60.0, uhi = 10.0, step = 1.0 / 16.0;
    const int n = (int)((uhi - ulo) / step);
    double best = -INFINITY;
    int ib = 0;
    for (int i = 0; i <= n; ++i) {
        const double v = log_integrand(t, lalpha, ulo + step * i);
        if (v > best) { best = v; ib = i; }
    }
    if (best == -INFINITY) return -INFINITY;
    int ia = ib, iz = ib;
    while (ia > 0 && log_integrand(t, lalpha, ulo + step * ia) > best - 60.0) --ia;
    while (iz < n && log_integrand(t, lalpha, ulo + step * iz) > best - 60.0) ++iz;
    const double a = ulo + step * ia, b = ulo + step * iz;
    const int panels = (int)ceil((b - a) * 32.0);
    const double h = (b - a) / panels;
    double acc = 0.0;
    for (int p = 0; p < panels; ++p) {
        const double mid = a + h * (p + 0.5), half = 0.5 * h;
        for (int q = 0; q < 8; ++q) {
            acc += GL16_W[q] * half * exp(log_integrand(t, lalpha, mid - half * GL16_X[q]) - best);
            acc += GL16_W[q] * half * exp(log_integrand(t, lalpha, mid + half * GL16_X[q]) - best);
        }
    }
    return best + log(acc);
}

static double log_pfa(const tstat *t, double lalpha) {
    if (t->mode == ATKDSP_CFAR_CA) {
        /* P(X > alpha T) = I_{1/(1+c)}(ne K, K), c = alpha / ne */
        const double c = exp(lalpha) / t->ne;
        double li, l1i;
        lbeta_i(t->ne * t->K, t->K, -log1p(c), log(c) - log1p(c), &li, &l1i);
        return li;
    }
    return log_integral(t, lalpha);
}

/* E[T] = integral of (1 - F_T(y)) dy, in u = ln y */
static double mean_t(const tstat *t) {
    const double a = -40.0, b = 8.0;
    const int panels = (int)((b - a) * 32.0);
    const double h = (b - a) / panels;
    double acc = 0.0;
    for (int p = 0; p < panels; ++p) {
        const double mid = a + h * (p + 0.5), half = 0.5 * h;
        for (int q = 0; q < 8; ++q) {
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                const double u = mid + sgn * half * GL16_X[q];
                double lf, lg;
                log_ft(t, exp(u), &lf, &lg);
                acc += GL16_W[q] * half * exp(lg + u);
            }
        }
    }
    return acc;
}

/* alpha with ln Pfa(alpha) = ln pfa, searched in [1e-3, 1e18] (the root must
 * lie inside, or the design is refused) */
static int solve(const tstat *t, double pfa, double *alpha) {
    const double target = log(pfa);
    double a = log(1e-3), b = log(1e18);
    double fa = log_pfa(t, a) - target, fb = log_pfa(t, b) - target;
    if (!(fa > 0.0 && fb < 0.0)) return ATKDSP_E_ARG;
    for (int i = 0; i < 40; ++i) {                           /* bisection to a tight bracket */
        const double m = 0.5 * (a + b), fm = log_pfa(t, m) - target;
        if (fm > 0.0) { a = m; fa = fm; } else { b = m; fb = fm; }
        if (b - a < 1e-6) break;
    }
    /* Brent's method on [a, b] */
    double c = a, fc = fa, d = b - a, e = d;
    for (int it = 0; it < 200; ++it) {
        if ((fb > 0.0) == (fc > 0.0)) { c = a; fc = fa; d = e = b - a; }
        if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
        const double tol = 2e-16 * fabs(b) + 1e-15, m = 0.5 * (c - b);
        if (fabs(m) <= tol || fb == 0.0) break;
        if (fabs(e) >= tol && fabs(fa) > fabs(fb)) {
            double p, q, r;
            const double s = fb / fa;
            if (a == c) {
                p = 2.0 * m * s;
                q = 1.0 - s;
            } else {
                q = fa / fc;
                r = fb / fc;
                p = s * (2.0 * m * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q; else p = -p;
            if (2.0 * p < fmin(3.0 * m * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
            else { d = m; e = m; }
        } else {
            d = m;
            e = m;
        }
        a = b;
        fa = fb;
        b += fabs(d) > tol ? d : (m > 0.0 ? tol : -tol);
        fb = log_pfa(t, b) - target;
    }
    *alpha = exp(b);
    return ATKDSP_OK;
}

static int rho2_valid(const double *rho2) {
    if (!rho2) return 1;
    for (int m = 0; m < CFAR_RHO_MAX; ++m)
        if (!(rho2[m] >= 0.0 && rho2[m] < 1.0)) return 0;
    return 1;
}

int atkdsp_cfar_alpha(int mode, int n_ref, double pfa, double looks, int os_k, const double *rho2, double *alpha,
                      double *et) {
    if (mode < ATKDSP_CFAR_CA || mode > ATKDSP_CFAR_SO || n_ref < 1 || n_ref > 256 || !(pfa >= 1e-15 && pfa <= 0.5) ||
        !(looks >= 0.25 && looks <= 64e6) || !alpha || !rho2_valid(rho2))
        return ATKDSP_E_ARG;
    const int N = 2 * n_ref;
    const int k = os_k ? os_k : (3 * N + 3) / 4;
    if (mode == ATKDSP_CFAR_OS && (k < 1 || k > N)) return ATKDSP_E_ARG;
    const double ne = atkdsp_cfar_n_eff(n_ref, rho2);
    tstat t = {mode, N, k, ne, looks};
    double target = pfa;
    if (mode == ATKDSP_CFAR_OS && ne != (double)N) {
        /* the rate the independent-cell CA test has at the correlated CA
         * threshold; OS is then solved at that rate with its integer N */
        const tstat ca = {ATKDSP_CFAR_CA, N, 0, ne, looks};
        double a_ca;
        const int rc = solve(&ca, pfa, &a_ca);
        if (rc) return rc;
        const tstat ind = {ATKDSP_CFAR_CA, N, 0, (double)N, looks};
        target = exp(log_pfa(&ind, log(a_ca)));
        if (!(target > 0.0)) target = 1e-300;
        t.ne = (double)N;
    }
    const int rc = solve(&t, target, alpha);
    if (rc) return rc;
    if (et) *et = mean_t(&t);
    return ATKDSP_OK;
}
