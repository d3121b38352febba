// Register-resident Stockham FFT for power-of-two rows (gfx950).
//
// A team of L lanes (inside one wavefront, or several wavefronts with
// workgroup barriers) transforms one row of N complex values; each lane keeps V = N / L values in VGPRs for the whole
// transform.  Pass p (radix R, span NS) works on butterflies
// j = lane + L*b, b < N/(R*L): inputs x[j + r*N/R], twiddle w_N^(r*(j mod NS)
// * N/(NS*R)), radix-R DFT, outputs y[(j - j mod NS)*R + j mod NS + r*NS]
// (the Stockham autosort step of Govindaraju et al., SC'08).  Between passes
// the values go once through the team's LDS row (write all, read all); the
// first pass reads x[lane + L*b + r*N/R] and the last pass produces
// y[lane + L*b + r*N/R], both unit-stride across lanes, so they map to
// coalesced global loads/stores or conflict-free LDS accesses.
//
// LDS rows are padded by one element every 16 (pad_index) so that the
// stride-R writes after the first pass do not fall into one bank.
#pragma once

#include "blink_fft.hpp"

namespace blink {

__device__ __forceinline__ constexpr int pad_index(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int padded_len(int n) { return n + (n >> 4); }

template <int S, int R>
__device__ __forceinline__ void dft_r(float2* x) {
    if constexpr (R == 2) dft2<S>(x);
    else if constexpr (R == 4) dft4<S>(x);
    else if constexpr (R == 8) dft8<S>(x);
    else dft16<S>(x);
}

// Radix schedule of a power-of-two N for L lanes: as many radix-16 passes as
// possible, then one smaller radix; every pass needs (N / R) % L == 0.
struct RegRadices {
    int n_pass;
    int r[8];
};

__host__ __device__ constexpr RegRadices reg_radices(int n, int lanes) {
    RegRadices rr{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    int rem = n;
    while (rem > 1) {
        int r = 16;
        while (r > 1 && (rem % r != 0 || (n / r) % lanes != 0)) r /= 2;
        if (r < 2) return RegRadices{-1, {0, 0, 0, 0, 0, 0, 0, 0}};
        rr.r[rr.n_pass++] = r;
        rem /= r;
    }
    return rr;
}

__host__ __device__ constexpr int reg_span(RegRadices rr, int p) {
    int s = 1;
    for (int i = 0; i < p; i++) s *= rr.r[i];
    return s;
}

// Twiddle sources for RegFft::run: w(m) = exp(+2 pi i m / N), 0 <= m < N.
struct TwTable {  // a full table of N entries (LDS or global)
    const float2* t;
    __device__ __forceinline__ float2 operator()(int m) const { return t[m]; }
};
struct TwSplit64 {  // two 64-entry tables: w(m) = lo[m % 64] * hi[m / 64], N <= 4096
    const float2* lo;
    const float2* hi;
    __device__ __forceinline__ float2 operator()(int m) const {
        return cmul(lo[m & 63], hi[m >> 6]);
    }
};

// kBlock: the team spans several wavefronts (L > 64), so the LDS exchanges
// synchronise the whole workgroup, and every team of the workgroup must run
// the same passes.
template <int N, int L, bool kBlock = (L > 64)>
struct RegFft {
    static constexpr int V = N / L;
    static constexpr RegRadices RAD = reg_radices(N, L);
    static_assert(RAD.n_pass > 0, "unsupported size / lane count");

    // Load x[lane + L*b + r*N/R0] (first-pass input order) through `ld(i)`.
    template <class Load>
    __device__ __forceinline__ static void load(float2 (&v)[V], int lane, Load ld) {
        constexpr int R = RAD.r[0];
        constexpr int NB = N / R, B = NB / L;
#pragma unroll
        for (int b = 0; b < B; b++)
#pragma unroll
            for (int r = 0; r < R; r++) v[b * R + r] = ld(lane + L * b + r * NB);
    }

    // Store y[lane + L*b + r*N/R_last] (last-pass output order) through `st(i, v)`.
    template <class Store>
    __device__ __forceinline__ static void store(const float2 (&v)[V], int lane, Store st) {
        constexpr int R = RAD.r[RAD.n_pass - 1];
        constexpr int NB = N / R, B = NB / L;
#pragma unroll
        for (int b = 0; b < B; b++)
#pragma unroll
            for (int r = 0; r < R; r++) st(lane + L * b + r * NB, v[b * R + r]);
    }

    template <int S, int P, class TW>
    __device__ __forceinline__ static void pass(float2 (&v)[V], int lane, const TW& tw) {
        constexpr int R = RAD.r[P];
        constexpr int NS = reg_span(RAD, P);
        constexpr int NB = N / R, B = NB / L;
#pragma unroll
        for (int b = 0; b < B; b++) {
            if constexpr (NS > 1) {
                const int j = lane + L * b;
                const int k = j % NS;
                constexpr int step = N / (NS * R);
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw(r * k * step);
                    if (S < 0) w.y = -w.y;
                    v[b * R + r] = cmul(v[b * R + r], w);
                }
            }
            dft_r<S, R>(&v[b * R]);
        }
    }

    // Stockham output order of pass P -> input order of pass P+1, through the
    // team's LDS row (padded).  A team inside one wavefront executes its LDS
    // operations in order (the fences stop compiler reordering); a wider team
    // synchronises the workgroup.
    template <int P>
    __device__ __forceinline__ static void exchange(float2 (&v)[V], float2* row, int lane) {
        constexpr int R1 = RAD.r[P];
        constexpr int NS1 = reg_span(RAD, P);
        constexpr int NB1 = N / R1, B1 = NB1 / L;
        constexpr int R2 = RAD.r[P + 1];
        constexpr int NB2 = N / R2, B2 = NB2 / L;
#pragma unroll
        for (int b = 0; b < B1; b++) {
            const int j = lane + L * b;
            const int k = j % NS1;
            const int base = (j - k) * R1 + k;
#pragma unroll
            for (int r = 0; r < R1; r++) row[pad_index(base + r * NS1)] = v[b * R1 + r];
        }
        team_sync<kBlock>();
#pragma unroll
        for (int b = 0; b < B2; b++)
#pragma unroll
            for (int r = 0; r < R2; r++) v[b * R2 + r] = row[pad_index(lane + L * b + r * NB2)];
        team_sync<kBlock>();
    }

    template <int S, int P = 0, class TW>
    __device__ __forceinline__ static void run(float2 (&v)[V], float2* row, int lane,
                                               const TW& tw) {
        pass<S, P>(v, lane, tw);
        if constexpr (P + 1 < RAD.n_pass) {
            exchange<P>(v, row, lane);
            run<S, P + 1>(v, row, lane, tw);
        }
    }

};

}  // namespace blink
