// Radix butterflies and the per-thread stage logic of the LDS FFTs.
//
// Everything here is __host__ __device__ so that the same code runs on gfx950
// and inside the CPU thread-team emulation in tests/ (which checks the index
// arithmetic against a textbook DFT without a GPU).
//
// Conventions (forward transform, unscaled, e^{-2 pi i jk/n}; the reference
// uses the same sign and scaling, src/fft.c:105-106 and include/atkdsp.h:379-383):
//   - a twiddle table tw[t] = exp(-2 pi i t / n), t in [0, n), is computed on
//     the host in FP64 and rounded to float once;
//   - a Stockham autosort stage of radix R with Ns = product of the earlier
//     radices maps butterfly j (k = j mod Ns) from x[j + r*n/R] to
//     y[(j/Ns)*Ns*R + k + m*Ns], after multiplying input r by tw[r*k*n/(Ns*R)].
#pragma once
#include <hip/hip_runtime.h>

#define FFT_HD __host__ __device__ __forceinline__

struct cpx { float x, y; };

FFT_HD cpx cmake(float a, float b) { cpx c; c.x = a; c.y = b; return c; }
FFT_HD cpx cadd(cpx a, cpx b) { return cmake(a.x + b.x, a.y + b.y); }
FFT_HD cpx csub(cpx a, cpx b) { return cmake(a.x - b.x, a.y - b.y); }
FFT_HD cpx cmul(cpx a, cpx b) { return cmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// multiply by -i (forward W4^1)
FFT_HD cpx cmul_mi(cpx a) { return cmake(a.y, -a.x); }
// multiply by +i
FFT_HD cpx cmul_pi(cpx a) { return cmake(-a.y, a.x); }
FFT_HD cpx cconj(cpx a) { return cmake(a.x, -a.y); }

// ---- forward radix kernels, in place on v[0..R) ----------------------------
FFT_HD void dft2(cpx &a, cpx &b) {
    cpx t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

// X0 = a+b+c+d, X1 = a-ib-c+id, X2 = a-b+c-d, X3 = a+ib-c-id
FFT_HD void dft4(cpx &a, cpx &b, cpx &c, cpx &d) {
    cpx s0 = cadd(a, c), d0 = csub(a, c);
    cpx s1 = cadd(b, d), d1 = cmul_mi(csub(b, d));
    a = cadd(s0, s1);
    c = csub(s0, s1);
    b = cadd(d0, d1);
    d = csub(d0, d1);
}

// exp(-2 pi i m/8) for m = 1, 3 (m = 2 is -i)
#define FFT_R8 0.70710678118654752440f

FFT_HD void dft8(cpx *v) {
    // n = 2*n1 + n2, k = k1 + 4*k2: DFT4 over n1 for each n2, twiddle W8^{n2 k1}, DFT2 over n2
    cpx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    // twiddles W8^k1 on the odd branch: 1, (1-i)/sqrt2, -i, (-1-i)/sqrt2
    o1 = cmake((o1.x + o1.y) * FFT_R8, (o1.y - o1.x) * FFT_R8);
    o2 = cmul_mi(o2);
    o3 = cmake((o3.y - o3.x) * FFT_R8, -(o3.x + o3.y) * FFT_R8);
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

// cos/sin(2 pi m/16) for the W16 twiddles
#define FFT_C1 0.92387953251128675613f
#define FFT_S1 0.38268343236508977173f

FFT_HD void dft16(cpx *v) {
    // n = 4*n1 + n2, k = k1 + 4*k2
    cpx a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        a[n2][0] = v[n2]; a[n2][1] = v[4 + n2]; a[n2][2] = v[8 + n2]; a[n2][3] = v[12 + n2];
        dft4(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
    }
    // twiddle W16^{n2*k1}, W16^m = cos(2 pi m/16) - i sin(2 pi m/16)
    // n2=1: m=1,2,3 ; n2=2: m=2,4,6 ; n2=3: m=3,6,9
    const cpx w1 = cmake(FFT_C1, -FFT_S1);
    const cpx w2 = cmake(FFT_R8, -FFT_R8);
    const cpx w3 = cmake(FFT_S1, -FFT_C1);
    const cpx w6 = cmake(-FFT_R8, -FFT_R8);
    const cpx w9 = cmake(-FFT_C1, FFT_S1);
    a[1][1] = cmul(a[1][1], w1); a[1][2] = cmul(a[1][2], w2); a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2); a[2][2] = cmul_mi(a[2][2]); a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3); a[3][2] = cmul(a[3][2], w6); a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cpx b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
        dft4(b0, b1, b2, b3);
        v[k1] = b0; v[k1 + 4] = b1; v[k1 + 8] = b2; v[k1 + 12] = b3;
    }
}

template <int R>
FFT_HD void dft_pow2(cpx *v) {
    if constexpr (R == 2) dft2(v[0], v[1]);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 16) dft16(v);
}

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// ---- power-of-two team FFT ---------------------------------------------------
// A team of NT = N/PT threads transforms N points. Thread j holds x[j + r*NT]
// (r < PT) on entry and X[j + r*NT] on exit. Full stages have radix PT; a
// final stage of radix RL = N / PT^S (RL < PT) runs PT/RL butterflies a thread.
//
// The team's LDS holds element i at lpad(i) = i + i/16 (N + N/16 slots): the
// first stage writes at a stride of PT elements, which without the pad would
// put every lane of a wavefront on the same LDS bank.
FFT_HD int lpad(int i) { return i + (i >> 4); }
template <int N>
constexpr int lpad_size() { return N + N / 16 + 1; }
template <int N, int PT>
struct TeamFftPlan {
    static constexpr int NT = N / PT;
    static constexpr int S = ilog2c(N) / ilog2c(PT);          // full stages
    static constexpr int RL = N / ipow(PT, S);                 // last radix (1 = none)
    static constexpr int NSTAGE = S + (RL > 1 ? 1 : 0);
    static_assert(ipow(PT, S) * RL == N, "N must be a power of two");
};

// One stage for thread j: v holds x[j + r*NT]. Writes the stage output to
// lds at Stockham positions, or, if `last`, leaves X[j + r*NT] in v.
template <int N, int PT, int R>
FFT_HD void team_stage(cpx *v, cpx *lds, const cpx *tw, int j, int Ns, bool last) {
    constexpr int NT = N / PT;
    constexpr int NB = PT / R;          // butterflies per thread
    cpx out[PT];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int jj = j + b * NT;
        const int k = jj % Ns;
        cpx u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = v[b + r * NB];
        if (Ns > 1 && k != 0) {
            const int step = k * (N / (Ns * R));
#pragma unroll
            for (int r = 1; r < R; ++r) u[r] = cmul(u[r], tw[r * step]);
        }
        dft_pow2<R>(u);
        if (last) {
#pragma unroll
            for (int m = 0; m < R; ++m) out[b + m * NB] = u[m];
        } else {
            const int base = (jj / Ns) * Ns * R + k;
#pragma unroll
            for (int m = 0; m < R; ++m) lds[lpad(base + m * Ns)] = u[m];
        }
    }
    if (last) {
#pragma unroll
        for (int r = 0; r < PT; ++r) v[r] = out[r];
    }
}

// The same full-radix stage (one butterfly a thread, N = PT^S) with the
// thread's twiddles for this stage given in w[r-1] = tw[r * step]: a thread
// transforming many frames holds them in registers. Value for value the
// arithmetic of team_stage<N, PT, PT>.
template <int N, int PT>
FFT_HD void team_stage_twiddles(cpx *w, const cpx *tw, int j, int Ns) {
    const int k = j % Ns;
    const int step = k * (N / (Ns * PT));
    for (int r = 1; r < PT; ++r) w[r - 1] = tw[r * step];
}

template <int N, int PT>
FFT_HD void team_stage_w(cpx *v, cpx *lds, const cpx *w, int j, int Ns, bool last) {
    const int k = j % Ns;
    if (Ns > 1 && k != 0) {
#pragma unroll
        for (int r = 1; r < PT; ++r) v[r] = cmul(v[r], w[r - 1]);
    }
    dft_pow2<PT>(v);
    if (!last) {
        const int base = (j / Ns) * Ns * PT + k;
#pragma unroll
        for (int m = 0; m < PT; ++m) lds[lpad(base + m * Ns)] = v[m];
    }
}

template <int N, int PT>
FFT_HD void team_load(cpx *v, const cpx *lds, int j) {
    constexpr int NT = N / PT;
    if constexpr (NT % 16 == 0) {
        // lpad(j + r*NT) = lpad(j) + r*(NT + NT/16) for j >= 0: one address,
        // the rest immediate offsets
        const cpx *b = lds + lpad(j);
#pragma unroll
        for (int r = 0; r < PT; ++r) v[r] = b[r * (NT + NT / 16)];
    } else {
#pragma unroll
        for (int r = 0; r < PT; ++r) v[r] = lds[lpad(j + r * NT)];
    }
}

// ---- generic mixed-radix Stockham stage (runtime radix) ----------------------
// Thread-per-output form: output o of stage (Ns, R) is y_m of butterfly j with
// o = (j/Ns)*Ns*R + k + m*Ns. Each output sums R twiddled inputs, so a large
// prime radix spreads over all threads instead of one.
FFT_HD cpx generic_stage_output(const cpx *in, const cpx *tw, int n, int Ns, int R, int o) {
    const int k = o % Ns;
    const int m = (o / Ns) % R;
    const int g = o / (Ns * R);
    const int j = g * Ns + k;
    const int nr = n / R;
    const int step = k * (n / (Ns * R));     // input twiddle index unit
    // y_m = sum_r x[j + r nr] * W_n^{r*step} * W_R^{r m}
    //     = sum_r x[j + r nr] * W_n^{r*(step + m*nr)}   (W_R = W_n^{nr})
    const long long unit = (long long)step + (long long)m * nr;
    const long long nn = n;
    long long e = 0;
    if (R <= 16) {
        float ax = 0.f, ay = 0.f;
        for (int r = 0; r < R; ++r) {
            const cpx a = in[j + r * nr];
            const cpx w = tw[(int)e];
            ax += a.x * w.x - a.y * w.y;
            ay += a.x * w.y + a.y * w.x;
            e += unit;
            if (e >= nn) e -= nn * (e / nn);
        }
        return cmake(ax, ay);
    }
    // a long prime radix sums many terms: accumulate in FP64 so the error
    // stays at the twiddles' own rounding
    double ax = 0.0, ay = 0.0;
    for (int r = 0; r < R; ++r) {
        const cpx a = in[j + r * nr];
        const cpx w = tw[(int)e];
        ax += (double)a.x * w.x - (double)a.y * w.y;
        ay += (double)a.x * w.y + (double)a.y * w.x;
        e += unit;
        if (e >= nn) e -= nn * (e / nn);
    }
    return cmake((float)ax, (float)ay);
}
