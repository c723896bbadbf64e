// CPU emulation of the power-of-two team FFT (atkdsp_amd/csrc/dev/fft_core.hpp):
// every thread of a team runs each stage in turn, with the LDS as an array and
// the barriers as the boundaries between loops. Prints the max relative error
// against an FP64 DFT for each size; the test asserts on it.
#include <cmath>
#include <cstdio>
#include <vector>
#include "fft_core.hpp"

template <int N, int PT>
static double run_one() {
    constexpr int NT = N / PT;
    using P = TeamFftPlan<N, PT>;
    std::vector<cpx> x(N), lds(lpad_size<N>());
    std::vector<cpx> tw(N);
    for (int t = 0; t < N; ++t) {
        const double a = 2.0 * M_PI * t / N;
        tw[t] = cmake((float)std::cos(a), (float)-std::sin(a));
    }
    unsigned s = 12345;
    for (int i = 0; i < N; ++i) {
        s = s * 1103515245u + 12345u;
        const float a = (float)((s >> 8) & 0xffff) / 65536.0f - 0.5f;
        s = s * 1103515245u + 12345u;
        const float b = (float)((s >> 8) & 0xffff) / 65536.0f - 0.5f;
        x[i] = cmake(a, b);
    }
    std::vector<std::vector<cpx>> v(NT, std::vector<cpx>(PT));
    for (int j = 0; j < NT; ++j)
        for (int r = 0; r < PT; ++r) v[j][r] = x[j + r * NT];
    int Ns = 1;
    for (int st = 0; st < P::NSTAGE; ++st) {
        const bool last = st == P::NSTAGE - 1;
        const bool full = st < P::S;
        for (int j = 0; j < NT; ++j) {
            if (full) team_stage<N, PT, PT>(v[j].data(), lds.data(), tw.data(), j, Ns, last);
            else team_stage<N, PT, P::RL>(v[j].data(), lds.data(), tw.data(), j, Ns, true);
        }
        if (!last)
            for (int j = 0; j < NT; ++j) team_load<N, PT>(v[j].data(), lds.data(), j);
        Ns *= full ? PT : P::RL;
    }
    double num = 0, den = 0;
    for (int k = 0; k < N; ++k) {
        double re = 0, im = 0;
        for (int n = 0; n < N; ++n) {
            const double a = -2.0 * M_PI * (double)((long long)n * k % N) / N;
            re += x[n].x * std::cos(a) - x[n].y * std::sin(a);
            im += x[n].x * std::sin(a) + x[n].y * std::cos(a);
        }
        const cpx g = v[k % NT][k / NT];
        num = std::fmax(num, std::hypot(g.x - re, g.y - im));
        den = std::fmax(den, std::hypot(re, im));
    }
    return num / den;
}

// The full-radix stages with each thread's twiddles precomputed (the
// display kernel's register form): must equal team_stage's result bit for bit.
template <int N, int PT>
static int run_tw_equal() {
    constexpr int NT = N / PT;
    using P = TeamFftPlan<N, PT>;
    static_assert(P::RL == 1, "full radix");
    std::vector<cpx> x(N), lds(lpad_size<N>()), lds2(lpad_size<N>()), tw(N);
    for (int t = 0; t < N; ++t) {
        const double a = 2.0 * M_PI * t / N;
        tw[t] = cmake((float)std::cos(a), (float)-std::sin(a));
    }
    unsigned s = 777;
    for (int i = 0; i < N; ++i) {
        s = s * 1103515245u + 12345u;
        const float a = (float)((s >> 8) & 0xffff) / 65536.0f - 0.5f;
        s = s * 1103515245u + 12345u;
        x[i] = cmake(a, (float)((s >> 8) & 0xffff) / 65536.0f - 0.5f);
    }
    std::vector<std::vector<cpx>> v(NT, std::vector<cpx>(PT)), u(NT, std::vector<cpx>(PT));
    std::vector<std::vector<cpx>> w(NT, std::vector<cpx>((P::S) * (PT - 1)));
    for (int j = 0; j < NT; ++j) {
        int Ns = PT;
        for (int st = 1; st < P::S; ++st, Ns *= PT) team_stage_twiddles<N, PT>(&w[j][(st - 1) * (PT - 1)], tw.data(), j, Ns);
        for (int r = 0; r < PT; ++r) v[j][r] = u[j][r] = x[j + r * NT];
    }
    int Ns = 1;
    for (int st = 0; st < P::S; ++st, Ns *= PT) {
        const bool last = st == P::S - 1;
        for (int j = 0; j < NT; ++j) {
            team_stage<N, PT, PT>(v[j].data(), lds.data(), tw.data(), j, Ns, last);
            team_stage_w<N, PT>(u[j].data(), lds2.data(), st ? &w[j][(st - 1) * (PT - 1)] : nullptr, j, Ns, last);
        }
        if (!last)
            for (int j = 0; j < NT; ++j) {
                team_load<N, PT>(v[j].data(), lds.data(), j);
                team_load<N, PT>(u[j].data(), lds2.data(), j);
            }
    }
    for (int j = 0; j < NT; ++j)
        for (int r = 0; r < PT; ++r)
            if (v[j][r].x != u[j][r].x || v[j][r].y != u[j][r].y) return 1;
    return 0;
}

int main() {
    std::printf("64 %.3e\n", run_one<64, 8>());
    std::printf("128 %.3e\n", run_one<128, 8>());
    std::printf("256 %.3e\n", run_one<256, 16>());
    std::printf("512 %.3e\n", run_one<512, 16>());
    std::printf("1024 %.3e\n", run_one<1024, 16>());
    std::printf("2048 %.3e\n", run_one<2048, 16>());
    std::printf("4096 %.3e\n", run_one<4096, 16>());
    std::printf("tw4096 %d\n", run_tw_equal<4096, 16>());
    std::printf("tw256 %d\n", run_tw_equal<256, 16>());
    return 0;
}
