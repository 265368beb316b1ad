// xc_fft.hpp -- LDS-staged Stockham FFT building blocks for CDNA4 (gfx950).
//
// One length-N complex transform is carried by TPF = N/E threads, each
// holding E = min(8, N) elements in registers in the "canonical" layout
//     reg[m] = x[t + TPF*m],  m in [0, E)
// Passes are radix-8 with one final radix-4/2 pass.  A pass reads its R
// inputs j + r*N/R from registers (they are exactly reg[b + r*E/R] for
// butterfly j = t + TPF*b), multiplies by the Stockham twiddles, runs the
// radix-R butterfly, and scatters the outputs to LDS at
//     (j/Ns)*Ns*R + j%Ns + r*Ns
// from which the next pass reloads the canonical layout.  The last pass's
// outputs land back in canonical order in registers, so a forward and an
// inverse transform chain without touching LDS (used to keep the cross
// power in registers between the column forward and inverse transforms).
//
// No rocFFT, no MFMA: an FFT is not a dense contraction.  Twiddles come from
// one host-generated FP64-accurate table of e^{-2 pi i m / XC_TW_N}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace xc {

constexpr int XC_TW_N = 8192;          // twiddle table length (largest Nx)
constexpr int XC_MAX_NH = 4096;        // largest complex transform length

// Complex value; aligned to its full size so that LDS and global accesses move
// it as one 8-byte (FP32) or 16-byte (FP64) access.
template <typename T> struct alignas(2 * sizeof(T)) cpx { T x, y; };

template <typename T> __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }
template <typename T> __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }

// Multiply by -i (DIR < 0, forward) or +i (DIR > 0, inverse).
template <int DIR, typename T> __device__ __forceinline__ cpx<T> rot90(cpx<T> a)
{
    if constexpr (DIR < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}

template <int DIR, typename T> __device__ __forceinline__ void bfly2(cpx<T> *v)
{
    cpx<T> t = v[0] - v[1];
    v[0] = v[0] + v[1];
    v[1] = t;
}

template <int DIR, typename T> __device__ __forceinline__ void bfly4(cpx<T> *v)
{
    cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
    cpx<T> t2 = v[1] + v[3], t3 = rot90<DIR>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
}

template <int DIR, typename T> __device__ __forceinline__ void bfly8(cpx<T> *v)
{
    const T s = (T)0.70710678118654752440;
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    bfly4<DIR>(e);
    bfly4<DIR>(o);
    if constexpr (DIR < 0) {
        o[1] = {s * (o[1].x + o[1].y), s * (o[1].y - o[1].x)};
        o[2] = {o[2].y, -o[2].x};
        o[3] = {s * (o[3].y - o[3].x), -s * (o[3].x + o[3].y)};
    } else {
        o[1] = {s * (o[1].x - o[1].y), s * (o[1].x + o[1].y)};
        o[2] = {-o[2].y, o[2].x};
        o[3] = {-s * (o[3].x + o[3].y), s * (o[3].x - o[3].y)};
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = e[k] + o[k];
        v[k + 4] = e[k] - o[k];
    }
}

template <int DIR, int R, typename T> __device__ __forceinline__ void bfly(cpx<T> *v)
{
    if constexpr (R == 8) bfly8<DIR>(v);
    else if constexpr (R == 4) bfly4<DIR>(v);
    else if constexpr (R == 2) bfly2<DIR>(v);
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <int N> struct Plan {
    static constexpr int L = ilog2c(N);
    static constexpr int E = N >= 8 ? 8 : N;
    static constexpr int TPF = N / E;
    static constexpr int NP = L == 0 ? 0 : (L + 2) / 3;
    static constexpr int radix(int p) { return (p < L / 3) ? 8 : (1 << (L % 3)); }
    static constexpr int stride_before(int p)
    {
        int s = 1;
        for (int q = 0; q < p; ++q) s *= radix(q);
        return s;
    }
};

// Twiddle e^{DIR 2 pi i m / n} for n a power of two dividing XC_TW_N.
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T> *__restrict__ tw, int m, int n)
{
    cpx<T> w = tw[(m & (n - 1)) * (XC_TW_N / n)];
    if constexpr (DIR > 0) w.y = -w.y;
    return w;
}

// Device memory through global-address-space pointers.  The kernels take
// their buffers out of descriptors as generic pointers, and a generic (flat)
// access also counts as an LDS operation, so every LDS wait would wait for it
// too; global loads and stores stay out of the LDS count.  Complex values move
// as two-element vectors (one 8-byte access for FP32).
template <typename T> struct Vec2;
template <> struct Vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct Vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };

template <typename T> __device__ __forceinline__ cpx<T> gld(const cpx<T> *p, long long i)
{
    typedef __attribute__((address_space(1))) const typename Vec2<T>::type gv;
    const typename Vec2<T>::type v = ((gv *)(const void *)p)[i];
    return cpx<T>{v.x, v.y};
}

template <typename T> __device__ __forceinline__ void gst(cpx<T> *p, long long i, cpx<T> c)
{
    typedef __attribute__((address_space(1))) typename Vec2<T>::type gv;
    typename Vec2<T>::type v;
    v.x = c.x;
    v.y = c.y;
    ((gv *)(void *)p)[i] = v;
}

// Scalar element i of global memory at p (any arithmetic type).
template <typename S> __device__ __forceinline__ S gld(const S *p, long long i)
{
    typedef __attribute__((address_space(1))) const S gs;
    return ((gs *)(const void *)p)[i];
}

// Scalar element i of global memory at p set to v.
template <typename S> __device__ __forceinline__ void gst(S *p, long long i, S v)
{
    typedef __attribute__((address_space(1))) S gs;
    ((gs *)(void *)p)[i] = v;
}

// Elements i and i + 1 of global memory at p set to a and b in one access.
template <typename S> __device__ __forceinline__ void gst2(S *p, long long i, S a, S b)
{
    typedef S v2 __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) v2 gv;
    v2 v;
    v.x = a;
    v.y = b;
    *((gv *)(void *)(p + i)) = v;
}

// 16 bytes of global memory at p (16-byte aligned).
__device__ __forceinline__ uint4 gld16(const void *p)
{
    typedef unsigned v4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) const v4 gv;
    const v4 v = *((gv *)p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Workgroup barrier that orders LDS only: this wave's LDS accesses complete,
// then s_barrier.  Unlike __syncthreads() it does not wait for the wave's
// global loads and stores, so table loads and row stores issued before it
// stay in flight across it.  Only for barriers whose purpose is an LDS
// exchange between the workgroup's threads.
__device__ __forceinline__ void lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// LDS view: element pos of this transform lives at base[(pos + pos/8) *
// stride].  The one-in-eight padding spreads the radix-8 scatter
// (pos = 8j + r across lanes j) over all banks instead of two.
template <typename T> struct LdsView {
    cpx<T> *base;
    int stride;
    __device__ __forceinline__ cpx<T> &operator[](int pos) const { return base[(pos + (pos >> 3)) * stride]; }
};

// LDS elements one padded length-n transform occupies.
constexpr int lds_len(int n) { return n + n / 8 + 1; }

// Per-thread Stockham twiddles of every pass, loaded once from the table
// before any data arrives (they depend only on the thread index), so the
// table latency is paid once instead of once per pass.  Forward sign;
// the inverse transform conjugates on use.
template <int N> struct TwPlan {
    using PL = Plan<N>;
    static constexpr int count(int upto)
    {
        int c = 0;
        for (int p = 1; p < upto && p < PL::NP; ++p) c += (PL::E / PL::radix(p)) * (PL::radix(p) - 1);
        return c;
    }
    static constexpr int TOTAL = count(32) > 0 ? count(32) : 1;
};

template <int N, typename T> struct TwReg {
    cpx<T> w[TwPlan<N>::TOTAL];
};

template <int N, typename T, int P = 1>
__device__ __forceinline__ void tw_load_pass(TwReg<N, T> &R, int t, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<N>;
    if constexpr (P < PL::NP) {
        constexpr int E = PL::E, TPF = PL::TPF, Rdx = PL::radix(P), NB = E / Rdx;
        constexpr int Ns = PL::stride_before(P), OFF = TwPlan<N>::count(P);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int k = (t + TPF * b) & (Ns - 1);
#pragma unroll
            for (int r = 1; r < Rdx; ++r) R.w[OFF + b * (Rdx - 1) + (r - 1)] = twiddle<-1>(tw, r * k, Ns * Rdx);
        }
        tw_load_pass<N, T, P + 1>(R, t, tw);
    }
}

template <int N, typename T>
__device__ __forceinline__ void tw_load(TwReg<N, T> &R, int t, const cpx<T> *__restrict__ tw)
{
    tw_load_pass<N, T, 1>(R, t, tw);
}

// One length-N transform in the canonical register layout.  Every thread
// of the workgroup must call it (it contains workgroup barriers), including
// threads whose data is padding.
// Twiddle sources: registers preloaded by tw_load (TwReg), or the table
// read lazily inside each pass (TwLazy).
template <typename T> struct TwLazy {
    const cpx<T> *tab;
};

template <int DIR, int N, typename T, int OFF, int B, int RR, int Rdx, int Ns>
__device__ __forceinline__ cpx<T> tw_get(const TwReg<N, T> &TW, int)
{
    cpx<T> w = TW.w[OFF + B * (Rdx - 1) + (RR - 1)];
    if constexpr (DIR > 0) w.y = -w.y;
    return w;
}

template <int DIR, int N, typename T, int OFF, int B, int RR, int Rdx, int Ns>
__device__ __forceinline__ cpx<T> tw_get(const TwLazy<T> &TW, int t)
{
    constexpr int TPF = Plan<N>::TPF;
    const int k = (t + TPF * B) & (Ns - 1);
    return twiddle<DIR>(TW.tab, RR * k, Ns * Rdx);
}

template <int DIR, int N, typename T, int P = 0, typename TWS>
__device__ __forceinline__ void fft_pass(cpx<T> (&v)[Plan<N>::E], int t, LdsView<T> lds, const TWS &TW)
{
    using PL = Plan<N>;
    if constexpr (P < PL::NP) {
        constexpr int E = PL::E, TPF = PL::TPF;
        constexpr int R = PL::radix(P);
        constexpr int NB = E / R;
        constexpr int Ns = PL::stride_before(P);
        constexpr int OFF = TwPlan<N>::count(P);
        // the pass's twiddles first: table loads issued before the barrier arrive during it
        cpx<T> w[NB][R];
        if constexpr (P > 0) {
            static_for<0, NB>([&](auto bI) {
                constexpr int b = decltype(bI)::value;
                static_for<1, R>([&](auto rI) {
                    constexpr int r = decltype(rI)::value;
                    w[b][r] = tw_get<DIR, N, T, OFF, b, r, R, Ns>(TW, t);
                });
            });
            // reload canonical layout written by the previous pass
            lds_sync();
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = lds[t + TPF * m];
        }
        cpx<T> u[NB][R];
        static_for<0, NB>([&](auto bI) {
            constexpr int b = decltype(bI)::value;
            static_for<0, R>([&](auto rI) {
                constexpr int r = decltype(rI)::value;
                u[b][r] = v[b + r * NB];
                if constexpr (P > 0 && r > 0) u[b][r] = cmul(u[b][r], w[b][r]);
            });
            bfly<DIR, R>(u[b]);
        });
        if constexpr (P + 1 == PL::NP) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < R; ++r) v[b + r * NB] = u[b][r];
        } else {
            lds_sync();  // everyone finished reading the previous contents
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int j = t + TPF * b;
                const int k = j & (Ns - 1);
                const int base = (j / Ns) * Ns * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) lds[base + r * Ns] = u[b][r];
            }
            fft_pass<DIR, N, T, P + 1, TWS>(v, t, lds, TW);
        }
    }
}

template <int DIR, int N, typename T>
__device__ __forceinline__ void fft(cpx<T> (&v)[Plan<N>::E], int t, LdsView<T> lds, const TwReg<N, T> &TW)
{
    fft_pass<DIR, N, T, 0, TwReg<N, T>>(v, t, lds, TW);
}

// Table form: each pass reads its twiddles from the table when it needs them.
template <int DIR, int N, typename T>
__device__ __forceinline__ void fft(cpx<T> (&v)[Plan<N>::E], int t, LdsView<T> lds,
                                    const cpx<T> *__restrict__ tw)
{
    const TwLazy<T> TW{tw};
    fft_pass<DIR, N, T, 0, TwLazy<T>>(v, t, lds, TW);
}

// RR independent transforms per thread, interleaved pass by pass: the passes'
// barriers and twiddles are shared, and the RR butterflies of a pass give the
// scheduler independent work.  Each transform has its own LDS view.
template <int DIR, int N, typename T, int RR, int P = 0, typename TWS>
__device__ __forceinline__ void fft_pass_rows(cpx<T> (&v)[RR][Plan<N>::E], int t, const LdsView<T> (&lds)[RR],
                                              const TWS &TW)
{
    using PL = Plan<N>;
    if constexpr (P < PL::NP) {
        constexpr int E = PL::E, TPF = PL::TPF;
        constexpr int R = PL::radix(P);
        constexpr int NB = E / R;
        constexpr int Ns = PL::stride_before(P);
        constexpr int OFF = TwPlan<N>::count(P);
        cpx<T> w[NB][R];
        if constexpr (P > 0) {
            static_for<0, NB>([&](auto bI) {
                constexpr int b = decltype(bI)::value;
                static_for<1, R>([&](auto rI) {
                    constexpr int r = decltype(rI)::value;
                    w[b][r] = tw_get<DIR, N, T, OFF, b, r, R, Ns>(TW, t);
                });
            });
            lds_sync();
#pragma unroll
            for (int q = 0; q < RR; ++q)
#pragma unroll
                for (int m = 0; m < E; ++m) v[q][m] = lds[q][t + TPF * m];
        }
        cpx<T> u[RR][NB][R];
        static_for<0, NB>([&](auto bI) {
            constexpr int b = decltype(bI)::value;
            static_for<0, R>([&](auto rI) {
                constexpr int r = decltype(rI)::value;
                if constexpr (P > 0 && r > 0) {
#pragma unroll
                    for (int q = 0; q < RR; ++q) u[q][b][r] = cmul(v[q][b + r * NB], w[b][r]);
                } else {
#pragma unroll
                    for (int q = 0; q < RR; ++q) u[q][b][r] = v[q][b + r * NB];
                }
            });
#pragma unroll
            for (int q = 0; q < RR; ++q) bfly<DIR, R>(u[q][b]);
        });
        if constexpr (P + 1 == PL::NP) {
#pragma unroll
            for (int q = 0; q < RR; ++q)
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int r = 0; r < R; ++r) v[q][b + r * NB] = u[q][b][r];
        } else {
            lds_sync();  // everyone finished reading the previous contents
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int j = t + TPF * b;
                const int k = j & (Ns - 1);
                const int base = (j / Ns) * Ns * R + k;
#pragma unroll
                for (int q = 0; q < RR; ++q)
#pragma unroll
                    for (int r = 0; r < R; ++r) lds[q][base + r * Ns] = u[q][b][r];
            }
            fft_pass_rows<DIR, N, T, RR, P + 1, TWS>(v, t, lds, TW);
        }
    }
}

// RR transforms with the twiddle source TW (TwLazy or TwReg); UPPER_ZERO:
// registers m >= E/2 are zero padding (constants, so the first pass drops
// their terms).
template <int DIR, int N, typename T, int RR, bool UPPER_ZERO, typename TWS>
__device__ __forceinline__ void fft_rows_tw(cpx<T> (&v)[RR][Plan<N>::E], int t, const LdsView<T> (&lds)[RR],
                                            const TWS &TW)
{
    if constexpr (UPPER_ZERO) {
        constexpr int E = Plan<N>::E;
#pragma unroll
        for (int q = 0; q < RR; ++q)
#pragma unroll
            for (int m = E / 2; m < E; ++m) v[q][m] = cpx<T>{(T)0, (T)0};
    }
    fft_pass_rows<DIR, N, T, RR, 0, TWS>(v, t, lds, TW);
}

// RR transforms with twiddles read from the table in each pass.
template <int DIR, int N, typename T, int RR, bool UPPER_ZERO>
__device__ __forceinline__ void fft_rows(cpx<T> (&v)[RR][Plan<N>::E], int t, const LdsView<T> (&lds)[RR],
                                         const cpx<T> *__restrict__ tw)
{
    const TwLazy<T> TW{tw};
    fft_rows_tw<DIR, N, T, RR, UPPER_ZERO, TwLazy<T>>(v, t, lds, TW);
}

// Forward or inverse transform of data whose registers m >= E/2 are zero
// (the upper half of the input is padding).  The zeros are constants here,
// so the first pass's butterflies lose their dead terms.
template <int DIR, int N, typename T, typename TWA>
__device__ __forceinline__ void fft_upper_zero(cpx<T> (&v)[Plan<N>::E], int t, LdsView<T> lds, const TWA &tw)
{
    constexpr int E = Plan<N>::E;
#pragma unroll
    for (int m = E / 2; m < E; ++m) v[m] = cpx<T>{(T)0, (T)0};
    fft<DIR, N>(v, t, lds, tw);
}

}  // namespace xc
