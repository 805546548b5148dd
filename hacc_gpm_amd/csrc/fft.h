// fft.h — Stockham FFT building blocks for batched 1-D transforms held in registers
// and LDS.  A workgroup transforms COLS columns of length N at once.  Stage s of a
// plan with radices R_0..R_{S-1} reads element j + r*N/R of each column, applies the
// twiddle W_{L R}^{k r} (k = j mod L, L = R_0*...*R_{s-1}), a radix-R DFT in
// registers, and writes output r to position (j/L)*L*R + k + r*L (Stockham
// auto-sort, no bit reversal).  Stage 0 reads global memory directly and the last
// stage writes it directly (its outputs land in natural order j + r*N/R), so a plan
// of S stages makes S-1 round trips through LDS.
#pragma once
#include "common.h"

// Everything below has internal linkage: k_fft.hip is compiled once per transform length
// and once more with another radix order for 1024, so Plan<1024> and the host-side
// twiddle builders differ between objects and must not be merged by the linker.
namespace {

// ------------------------------------------------------------ plans ----
template <int N> struct Plan;
#define HGPM_PLAN(N_, ...)                                                        \
    template <> struct Plan<N_> {                                                 \
        static constexpr int S = sizeof((int[]){__VA_ARGS__}) / sizeof(int);       \
        static __host__ __device__ constexpr int R(int s) {                       \
            constexpr int r[] = {__VA_ARGS__};                                    \
            return r[s];                                                          \
        }                                                                         \
    };
HGPM_PLAN(16, 16)
HGPM_PLAN(24, 3, 8)
HGPM_PLAN(32, 4, 8)
HGPM_PLAN(48, 3, 16)
HGPM_PLAN(64, 8, 8)
HGPM_PLAN(96, 3, 4, 8)
HGPM_PLAN(128, 8, 16)
HGPM_PLAN(192, 3, 8, 8)
HGPM_PLAN(256, 16, 16)
HGPM_PLAN(384, 3, 8, 16)
HGPM_PLAN(512, 8, 8, 8)
HGPM_PLAN(768, 3, 16, 16)
#if defined(HGPM_FFT_PLAN_ALT)
HGPM_PLAN(1024, 8, 8, 16) // the alternative radix order the autotuner may pick
#else
HGPM_PLAN(1024, 16, 8, 8)
#endif
#undef HGPM_PLAN

template <int N> __host__ __device__ constexpr int plan_L(int s) {
    int l = 1;
    for (int i = 0; i < s; i++) l *= Plan<N>::R(i);
    return l;
}
// Inverse plan: radices reversed, so that the last forward stage and the first
// inverse stage hold the same elements in the same thread.
template <int N> struct RevPlan {
    static constexpr int S = Plan<N>::S;
    static __host__ __device__ constexpr int R(int s) { return Plan<N>::R(S - 1 - s); }
};
template <int N> __host__ __device__ constexpr int rplan_L(int s) {
    int l = 1;
    for (int i = 0; i < s; i++) l *= RevPlan<N>::R(i);
    return l;
}

// ------------------------------------------------------ unit roots ----
// cos(m*pi/16), m = 0..8
__device__ __forceinline__ double cos_pi16(int m) {
    switch (m) {
    case 0: return 1.0;
    case 1: return 0.98078528040323044913;
    case 2: return 0.92387953251128675613;
    case 3: return 0.83146961230254523708;
    case 4: return 0.70710678118654752440;
    case 5: return 0.55557023301960222474;
    case 6: return 0.38268343236508977173;
    case 7: return 0.19509032201612826785;
    default: return 0.0;
    }
}
// v * exp(DIR * 2 pi i k / R) for R | 32 (k folded at compile time after unrolling)
template <int R, int DIR, typename V> __device__ __forceinline__ V twc(int k, V v) {
    typedef decltype(v.x) T;
    int m = (32 / R) * k; // angle = m * pi / 16, 0 <= m < 32
    if (m == 0) return v;
    if (m == 8) return cmul_mi<DIR>(v);   // exp(DIR i pi/2) = DIR i
    if (m == 16) return cmake<V>(-v.x, -v.y);
    if (m == 24) return cmul_mi<-DIR>(v); // exp(DIR i 3pi/2) = -DIR i
    T c, s;
    if (m <= 8) { c = (T)cos_pi16(m); s = (T)cos_pi16(8 - m); }
    else if (m <= 16) { c = -(T)cos_pi16(16 - m); s = (T)cos_pi16(m - 8); }
    else if (m <= 24) { c = -(T)cos_pi16(m - 16); s = -(T)cos_pi16(24 - m); }
    else { c = (T)cos_pi16(32 - m); s = -(T)cos_pi16(m - 24); }
    if (DIR < 0) s = -s;
    return cmake<V>(v.x * c - v.y * s, v.x * s + v.y * c);
}

// ------------------------------------------------- radix butterflies ----
// In-register DFT of R points, natural order in and out. DIR=-1: e^{-i}, +1: e^{+i}.
template <int R, int DIR, typename V> struct DFT {
    // power of two >= 4: decimation in frequency into two half-size DFTs
    static __device__ __forceinline__ void run(V* x) {
        constexpr int H = R / 2;
        V a[H], b[H];
#pragma unroll
        for (int k = 0; k < H; k++) {
            a[k] = cadd(x[k], x[k + H]);
            b[k] = twc<R, DIR>(k, csub(x[k], x[k + H]));
        }
        DFT<H, DIR, V>::run(a);
        DFT<H, DIR, V>::run(b);
#pragma unroll
        for (int m = 0; m < H; m++) {
            x[2 * m] = a[m];
            x[2 * m + 1] = b[m];
        }
    }
};
template <int DIR, typename V> struct DFT<1, DIR, V> {
    static __device__ __forceinline__ void run(V*) {}
};
template <int DIR, typename V> struct DFT<2, DIR, V> {
    static __device__ __forceinline__ void run(V* x) {
        V a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    }
};
template <int DIR, typename V> struct DFT<3, DIR, V> {
    static __device__ __forceinline__ void run(V* x) {
        typedef decltype(x[0].x) T;
        const T h = (T)0.86602540378443864676; // sqrt(3)/2
        V a = x[0], s = cadd(x[1], x[2]), d = csub(x[1], x[2]);
        x[0] = cadd(a, s);
        V m = cmake<V>(a.x - (T)0.5 * s.x, a.y - (T)0.5 * s.y);
        // r = DIR * h * i * d
        V r = cmake<V>(-(T)DIR * h * d.y, (T)DIR * h * d.x);
        x[1] = cadd(m, r);
        x[2] = csub(m, r);
    }
};

// twiddle from the table tw[e] = exp(-2 pi i e / N) (forward); conjugated for DIR>0
template <int DIR, typename V> __device__ __forceinline__ V twmul(V v, V w) {
    if (DIR > 0) w.y = -w.y;
    return cmul(v, w);
}

// Multiply x[r] by w^r, r = 1..R-1, where w^r = t[r*L] (conjugated for DIR > 0).
// For R in {4, 8, 16} only w, w^2 and w^4 are loaded and the other powers are
// composed as w^(4a+b) = (w^4)^a w^b, at most three complex products deep, which cuts
// the table loads of a radix-16 butterfly from 15 to 3.
template <int R, int DIR, typename V> __device__ __forceinline__ void apply_twiddles(V* x, const V* t, int L) {
    if constexpr (R == 4 || R == 8 || R == 16) {
        auto ld = [&](int r) {
            V w = t[r * L];
            if (DIR > 0) w.y = -w.y;
            return w;
        };
        const V w1 = ld(1), w2 = ld(2);
        const V w3 = cmul(w1, w2);
        x[1] = cmul(x[1], w1);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], w3);
        if constexpr (R >= 8) {
            const V w4 = ld(4);
            V q = w4;
#pragma unroll
            for (int a = 1; a < R / 4; a++) {
                if (a > 1) q = cmul(q, w4);
                x[4 * a] = cmul(x[4 * a], q);
                x[4 * a + 1] = cmul(x[4 * a + 1], cmul(q, w1));
                x[4 * a + 2] = cmul(x[4 * a + 2], cmul(q, w2));
                x[4 * a + 3] = cmul(x[4 * a + 3], cmul(q, w3));
            }
        }
    } else {
#pragma unroll
        for (int r = 1; r < R; r++) x[r] = twmul<DIR>(x[r], t[r * L]);
    }
}

// ------------------------------------------------------- tile engine ----
template <int I> struct IntC { static constexpr int value = I; };
template <int I, int S, typename F> __device__ __forceinline__ void static_for_impl(F& f) {
    if constexpr (I < S) {
        f(IntC<I>{});
        static_for_impl<I + 1, S>(f);
    }
}
template <int S, typename F> __device__ __forceinline__ void static_for(F& f) { static_for_impl<0, S>(f); }

// Per-thread register block for a stage of radix R over N x COLS elements.
template <int N, int COLS, int NT, int R> struct StageShape {
    static constexpr int NBF = N / R;         // butterflies per column
    static constexpr int TOT = NBF * COLS;    // butterflies per tile
    static constexpr int NB = (TOT + NT - 1) / NT;
};

// LDS slot of element n of column c.  ROW: columns are rows of the tile (lane order
// follows n); otherwise lane order follows c.  One pad slot per 32 elements breaks
// the power-of-two strides of the Stockham writes across LDS banks.
template <int N, int COLS, bool ROW> __device__ __forceinline__ int lds_slot(int n, int c) {
    int lin = ROW ? c * N + n : n * COLS + c;
    return lin + (lin >> 5);
}
template <int N, int COLS> constexpr int lds_elems() { return N * COLS + (N * COLS) / 32 + 1; }

// Stage s of plan P uses the twiddle table at tw + s*N, entry r*L + k =
// exp(-2 pi i k r / (L R)) (k < L, r < R); consecutive k are adjacent so the lanes
// of a wave read contiguous entries.
template <int N, bool REV> __host__ __device__ constexpr int stage_R(int s) {
    return REV ? RevPlan<N>::R(s) : Plan<N>::R(s);
}
template <int N, bool REV> __host__ __device__ constexpr int stage_L(int s) {
    return REV ? rplan_L<N>(s) : plan_L<N>(s);
}

// Transform a tile of COLS columns of length N.  ld(b, r, n, c) returns element n
// of column c (b, r: the register slot it lands in, compile-time after unrolling);
// st(b, r, n, c, v) consumes transformed element n of column c.  Stage 0 calls ld,
// the last stage calls st; the stages in between exchange through lds.  With
// REV = true the radices run in reverse order, so that ld of a REV transform can
// take the register slots st of a forward transform produced.
template <typename V, int N, int COLS, int NT, int DIR, bool REV, bool ROW, typename LD, typename ST>
__device__ __forceinline__ void fft_tile(LD&& ld, ST&& st, V* lds, const V* tw) {
    constexpr int S = Plan<N>::S;
    const int tid = threadIdx.x;
    auto body = [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int R = stage_R<N, REV>(s);
        constexpr int L = stage_L<N, REV>(s);
        typedef StageShape<N, COLS, NT, R> SS;
        V x[SS::NB][R];
#pragma unroll
        for (int b = 0; b < SS::NB; b++) {
            int w = tid + b * NT;
            if (SS::TOT % NT == 0 || w < SS::TOT) {
                int c = ROW ? w / SS::NBF : w % COLS;
                int j = ROW ? w % SS::NBF : w / COLS;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    int n = j + r * SS::NBF;
                    if constexpr (s == 0) x[b][r] = ld(b, r, n, c);
                    else x[b][r] = lds[lds_slot<N, COLS, ROW>(n, c)];
                }
            }
        }
        if constexpr (s > 0) __syncthreads();
#pragma unroll
        for (int b = 0; b < SS::NB; b++) {
            int w = tid + b * NT;
            if (SS::TOT % NT == 0 || w < SS::TOT) {
                int c = ROW ? w / SS::NBF : w % COLS;
                int j = ROW ? w % SS::NBF : w / COLS;
                int k = j % L;
                if constexpr (L > 1) apply_twiddles<R, DIR>(x[b], tw + s * N + k, L);
                DFT<R, DIR, V>::run(x[b]);
                if constexpr (s == S - 1) {
#pragma unroll
                    for (int r = 0; r < R; r++) st(b, r, j + r * SS::NBF, c, x[b][r]);
                } else {
                    int base = (j / L) * L * R + k;
#pragma unroll
                    for (int r = 0; r < R; r++) lds[lds_slot<N, COLS, ROW>(base + r * L, c)] = x[b][r];
                }
            }
        }
        if constexpr (s < S - 1) __syncthreads();
    };
    static_for<S>(body);
}

// Registers a thread holds for the first stage of a REV transform (= last stage of
// the forward transform): NB x R elements.
template <int N, int COLS, int NT> struct LastShape {
    static constexpr int R = Plan<N>::R(Plan<N>::S - 1);
    typedef StageShape<N, COLS, NT, R> SS;
    static constexpr int NB = SS::NB;
};

// Host-side twiddle table for plan (N, REV): S*N complex doubles (re, im).
template <int N, bool REV> inline void make_twiddles(double* out) {
    constexpr int S = Plan<N>::S;
    for (int i = 0; i < 2 * S * N; i++) out[i] = 0.0;
    for (int s = 0; s < S; s++) {
        int R = stage_R<N, REV>(s), L = stage_L<N, REV>(s);
        for (int r = 0; r < R; r++)
            for (int k = 0; k < L; k++) {
                double ang = -2.0 * HGPM_PI * (double)(k * r) / (double)(L * R);
                out[2 * (s * N + r * L + k)] = cos(ang);
                out[2 * (s * N + r * L + k) + 1] = sin(ang);
            }
    }
}
} // namespace
