// fft.hpp -- FP32 complex FFT building blocks for gfx950: in-register DFTs of
// radix 2/4/8/16 and a Stockham autosort pass over a column tile in LDS.
//
// Conventions: forward transforms use exp(-2 pi i n k / N), inverse
// transforms exp(+2 pi i n k / N), both unnormalised.  Twiddles come from a
// table generated in FP64 and rounded once to FP32.  Complex products use
// explicit fmaf (the library is otherwise compiled without contraction).
#pragma once

#include "sw_common.hpp"

namespace sw {
namespace fft {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a)
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// multiply by exp(-+ 2 pi i m / 8) for m = 1, 3 (m = 2 is rot90)
template <bool INV, int M>
__device__ __forceinline__ float2 rot45(float2 a)
{
    constexpr float h = 0.70710678118654752440f;
    // exp(-i pi/4) = h(1 - i); exp(-3 i pi/4) = h(-1 - i); inverse: conjugates
    if constexpr(M == 1) {
        return INV ? make_float2(h * (a.x - a.y), h * (a.x + a.y))
                   : make_float2(h * (a.x + a.y), h * (a.y - a.x));
    } else {
        return INV ? make_float2(-h * (a.x + a.y), h * (a.x - a.y))
                   : make_float2(h * (a.y - a.x), -h * (a.x + a.y));
    }
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b)
{
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

// natural order in and out
template <bool INV>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3)
{
    const float2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
    const float2 t2 = cadd(v1, v3), t3 = rot90<INV>(csub(v1, v3));
    v0 = cadd(t0, t2);
    v2 = csub(t0, t2);
    v1 = cadd(t1, t3);
    v3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* v)
{
    // n = 2 n1 + n2: DFT4 over n1, twiddle W8^(n2 k1), DFT2 over n2
    dft4<INV>(v[0], v[2], v[4], v[6]);
    dft4<INV>(v[1], v[3], v[5], v[7]);
    v[3] = rot45<INV, 1>(v[3]);
    v[5] = rot90<INV>(v[5]);
    v[7] = rot45<INV, 3>(v[7]);
    float2 o[8];
#pragma unroll
    for(int k1 = 0; k1 < 4; ++k1) {
        o[k1]     = cadd(v[2 * k1], v[2 * k1 + 1]);
        o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
    }
#pragma unroll
    for(int k = 0; k < 8; ++k) v[k] = o[k];
}

// Butterflies with a twiddled second input in FMA form: a + w b, then
// a - w b formed as 2 a - (a + w b), 6 instructions instead of the 8 of a
// complex product followed by a sum and a difference.  The difference
// carries the rounding of the sum, an error of the same size as that of the
// product it replaces.
__device__ __forceinline__ void bfly_w(float2& a, float2& b, float2 w)
{
    const float2 s = make_float2(fmaf(w.x, b.x, fmaf(-w.y, b.y, a.x)), fmaf(w.x, b.y, fmaf(w.y, b.x, a.y)));
    b = make_float2(fmaf(2.0f, a.x, -s.x), fmaf(2.0f, a.y, -s.y));
    a = s;
}
// rot45<INV, M>(a) / h: the two sums of the 45-degree rotation, unscaled
template <bool INV, int M>
__device__ __forceinline__ float2 rot45_sums(float2 a)
{
    if constexpr(M == 1)
        return INV ? make_float2(a.x - a.y, a.x + a.y) : make_float2(a.x + a.y, a.y - a.x);
    else
        return INV ? make_float2(-(a.x + a.y), a.x - a.y) : make_float2(a.y - a.x, -(a.x + a.y));
}
// bfly_w with w = exp(-+ 2 pi i M / 8), M = 1, 3: the product is h times two sums
template <bool INV, int M>
__device__ __forceinline__ void bfly_w8(float2& a, float2& b)
{
    constexpr float h = 0.70710678118654752440f;
    const float2 u = rot45_sums<INV, M>(b);
    const float2 s = make_float2(fmaf(h, u.x, a.x), fmaf(h, u.y, a.y));
    b = make_float2(fmaf(2.0f, a.x, -s.x), fmaf(2.0f, a.y, -s.y));
    a = s;
}

// DFT4 of (a0, a1, a2, a3) whose first two butterflies' sums and
// differences are already formed: t0, t1 = a0 +- a2 and t2, t3r = a1 +- a3
template <bool INV>
__device__ __forceinline__ void dft4_tail(float2& v0, float2& v1, float2& v2, float2& v3, float2 t0, float2 t1,
                                          float2 t2, float2 t3r)
{
    const float2 t3 = rot90<INV>(t3r);
    v0 = cadd(t0, t2);
    v2 = csub(t0, t2);
    v1 = cadd(t1, t3);
    v3 = csub(t1, t3);
}

// DFT16 after its first layer of DFT4s: v[4 k1 + n2] holds a[k1][n2]; the
// second layer takes DFT4 over n2 of a[k1][n2] W16^(n2 k1), with the
// twiddles folded into the butterflies, and the 4x4 register block is
// transposed to natural order
template <bool INV>
__device__ __forceinline__ void dft16_second(float2* v)
{
    constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    constexpr float h  = 0.70710678118654752440f;
    const float sg = INV ? 1.0f : -1.0f;
    // W16^m = cos(2 pi m/16) -+ i sin(2 pi m/16)
    const float2 w1 = make_float2(c1, sg * s1), w3 = make_float2(s1, sg * c1), w9 = make_float2(-c1, -sg * s1);
    dft4<INV>(v[0], v[1], v[2], v[3]);
    {   // k1 = 1: twiddles W16^1, W16^2 = W8^1, W16^3
        float2 a0 = v[4], a2 = v[6], p = cmul(v[5], w1), a3 = v[7];
        bfly_w8<INV, 1>(a0, a2);
        bfly_w(p, a3, w3);
        dft4_tail<INV>(v[4], v[5], v[6], v[7], a0, a2, p, a3);
    }
    {   // k1 = 2: twiddles W16^2 = W8^1, W16^4 = -+i, W16^6 = W8^3
        const float2 r2 = rot90<INV>(v[10]);
        const float2 t0 = cadd(v[8], r2), t1 = csub(v[8], r2);
        const float2 u = rot45_sums<INV, 1>(v[9]);
        float2 p = make_float2(h * u.x, h * u.y), a3 = v[11];
        bfly_w8<INV, 3>(p, a3);
        dft4_tail<INV>(v[8], v[9], v[10], v[11], t0, t1, p, a3);
    }
    {   // k1 = 3: twiddles W16^3, W16^6 = W8^3, W16^9
        float2 a0 = v[12], a2 = v[14], p = cmul(v[13], w3), a3 = v[15];
        bfly_w8<INV, 3>(a0, a2);
        bfly_w(p, a3, w9);
        dft4_tail<INV>(v[12], v[13], v[14], v[15], a0, a2, p, a3);
    }
    // X[k1 + 4 k2] = v[4 k1 + k2]: transpose the 4x4 register block
    float2 o[16];
#pragma unroll
    for(int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for(int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
    for(int k = 0; k < 16; ++k) v[k] = o[k];
}

template <bool INV>
__device__ __forceinline__ void dft16(float2* v)
{
    // n = 4 n1 + n2: DFT4 over n1 for each n2, twiddle W16^(n2 k1), DFT4 over n2
#pragma unroll
    for(int n2 = 0; n2 < 4; ++n2) dft4<INV>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    // a[k1][n2] now sits in v[4 k1 + n2]
    dft16_second<INV>(v);
}

// DFT16 of v[i] m[i]: the products are folded into the first layer's
// butterflies (v and m are read, v is overwritten with the transform).
// ONE0: m[0] is 1 and is not read.
template <bool INV, bool ONE0 = false>
__device__ __forceinline__ void dft16_mul(float2* v, const float2* m)
{
#pragma unroll
    for(int n2 = 0; n2 < 4; ++n2) {
        float2 p = (ONE0 && n2 == 0) ? v[0] : cmul(v[n2], m[n2]), b = v[8 + n2];
        float2 q = cmul(v[4 + n2], m[4 + n2]), d = v[12 + n2];
        bfly_w(p, b, m[8 + n2]);
        bfly_w(q, d, m[12 + n2]);
        dft4_tail<INV>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2], p, b, q, d);
    }
    dft16_second<INV>(v);
}

// cos, sin of 2 pi k / 32, k < 16
__device__ constexpr float kW32[16][2] = {
    {1.00000000000000000000f, 0.00000000000000000000f},
    {0.98078528040323043058f, 0.19509032201612824808f},
    {0.92387953251128673848f, 0.38268343236508978178f},
    {0.83146961230254523567f, 0.55557023301960217765f},
    {0.70710678118654757274f, 0.70710678118654746172f},
    {0.55557023301960228867f, 0.83146961230254523567f},
    {0.38268343236508983729f, 0.92387953251128673848f},
    {0.19509032201612833135f, 0.98078528040323043058f},
    {0.00000000000000006123f, 1.00000000000000000000f},
    {-0.19509032201612819257f, 0.98078528040323043058f},
    {-0.38268343236508972627f, 0.92387953251128673848f},
    {-0.55557023301960195560f, 0.83146961230254545772f},
    {-0.70710678118654746172f, 0.70710678118654757274f},
    {-0.83146961230254534669f, 0.55557023301960217765f},
    {-0.92387953251128673848f, 0.38268343236508989280f},
    {-0.98078528040323043058f, 0.19509032201612860891f}};

// The radix-2 step of a DFT32 from the DFT16s E of the even and O of the
// odd samples: X[k] = E[k] + W32^k O[k], X[k + 16] = E[k] - W32^k O[k]
template <bool INV>
__device__ __forceinline__ void dft32_combine(float2* v, float2* e, float2* o)
{
#pragma unroll
    for(int k = 0; k < 16; ++k) {
        if(k == 0) {
            v[0]  = cadd(e[0], o[0]);
            v[16] = csub(e[0], o[0]);
            continue;
        } else if(k == 8) {
            const float2 t = rot90<INV>(o[8]);
            v[8]  = cadd(e[8], t);
            v[24] = csub(e[8], t);
            continue;
        } else if(k == 4) {
            bfly_w8<INV, 1>(e[4], o[4]);
        } else if(k == 12) {
            bfly_w8<INV, 3>(e[12], o[12]);
        } else {
            bfly_w(e[k], o[k], make_float2(kW32[k][0], INV ? kW32[k][1] : -kW32[k][1]));
        }
        v[k]      = e[k];
        v[k + 16] = o[k];
    }
}

// natural order in and out: radix-2 split into two DFT16 of the even and odd
// samples
template <bool INV>
__device__ __forceinline__ void dft32(float2* v)
{
    float2 e[16], o[16];
#pragma unroll
    for(int i = 0; i < 16; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    dft16<INV>(e);
    dft16<INV>(o);
    dft32_combine<INV>(v, e, o);
}

// DFT32 of v[i] m[i] (natural order), the products folded into the DFT16s'
// first butterflies; ONE0: m[0] is 1 and is not read
template <bool INV, bool ONE0 = false>
__device__ __forceinline__ void dft32_mul(float2* v, const float2* m)
{
    float2 e[16], o[16], me[16], mo[16];
#pragma unroll
    for(int i = 0; i < 16; ++i) {
        e[i]  = v[2 * i];
        o[i]  = v[2 * i + 1];
        me[i] = (ONE0 && i == 0) ? make_float2(1.0f, 0.0f) : m[2 * i];
        mo[i] = m[2 * i + 1];
    }
    dft16_mul<INV, ONE0>(e, me);
    dft16_mul<INV>(o, mo);
    dft32_combine<INV>(v, e, o);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v)
{
    if constexpr(R == 2) {
        dft2<INV>(v[0], v[1]);
    } else if constexpr(R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else if constexpr(R == 8) {
        dft8<INV>(v);
    } else if constexpr(R == 16) {
        dft16<INV>(v);
    } else {
        static_assert(R == 32, "radix");
        dft32<INV>(v);
    }
}

// DFT_R of v[k] m[k] with m[0] = 1 (not read): the products folded into
// the first butterflies for R = 16, 32
template <int R, bool INV>
__device__ __forceinline__ void dft_mul(float2* v, const float2* m)
{
    if constexpr(R == 32) {
        dft32_mul<INV, true>(v, m);
    } else if constexpr(R == 16) {
        dft16_mul<INV, true>(v, m);
    } else {
#pragma unroll
        for(int k = 1; k < R; ++k) v[k] = cmul(v[k], m[k]);
        dft<R, INV>(v);
    }
}

// LDS index with one float2 of padding every 16 (breaks power-of-two strides)
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int padded(int n) { return n + (n >> 4); }

// Twiddle source for the passes: exp(-2 pi i m / L) = A[m / LO] * B[m % LO]
// with two small tables (LO = min(128, L) entries and L / LO entries) held
// in LDS; both are FP64-generated values rounded once to FP32.
struct Twiddles {
    const float2* A;  // exp(-2 pi i (LO j) / L), j < L / LO
    const float2* B;  // exp(-2 pi i j / L),      j < LO
    int lo_bits;
    __device__ __forceinline__ float2 operator()(int m) const
    {
        return cmul(A[m >> lo_bits], B[m & ((1 << lo_bits) - 1)]);
    }
};

template <int L>
struct TwiddleLds {
    static constexpr int LO = L < 128 ? L : 128;
    static constexpr int HI = L / LO;
    static constexpr int SIZE = LO + HI;
};

// Fill the two LDS twiddle tables from the full table tw[m] = exp(-2 pi i m / L).
template <int L, int NT>
__device__ __forceinline__ Twiddles load_twiddles(float2* lds, const float2* __restrict__ tw)
{
    using TL = TwiddleLds<L>;
    for(int i = threadIdx.x; i < TL::LO; i += NT) lds[i] = tw[i];
    for(int i = threadIdx.x; i < TL::HI; i += NT) lds[TL::LO + i] = tw[i * TL::LO];
    return Twiddles{lds + TL::LO, lds, __builtin_ctz(TL::LO)};
}

// w[k] = w1^k for k = 1..R-1 by a product tree of depth <= 4
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 w1, float2* w)
{
    w[1] = w1;
    if constexpr(R > 2) {
        w[2] = cmul(w1, w1);
        w[3] = cmul(w[2], w1);
    }
    if constexpr(R > 4) {
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[4], w1);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr(R > 8) {
        w[8] = cmul(w[4], w[4]);
#pragma unroll
        for(int k = 1; k < 8; ++k) w[8 + k] = cmul(w[8], w[k]);
    }
    if constexpr(R > 16) {
        w[16] = cmul(w[8], w[8]);
#pragma unroll
        for(int k = 1; k < 16; ++k) w[16 + k] = cmul(w[16], w[k]);
    }
}

// One Stockham autosort pass of radix R over C columns of length L held in
// LDS (column c at offset c*padded(L)).  Ns = product of the radices already
// applied.  All threads of the block call it.
template <int R, bool INV, int L, int C, int NT>
__device__ __forceinline__ void stockham_pass(float2* buf, int Ns, const Twiddles& tw)
{
    constexpr int NB  = L / R;                 // butterflies per column
    constexpr int PER = (C * NB + NT - 1) / NT; // butterflies per thread
    const int tid = threadIdx.x;
    float2 v[PER][R];
#pragma unroll
    for(int q = 0; q < PER; ++q) {
        const int item = tid + q * NT;
        if(item < C * NB) {
            const int c = item / NB, j = item % NB;
            const float2* col = buf + c * padded(L);
#pragma unroll
            for(int k = 0; k < R; ++k) v[q][k] = col[pad(j + k * NB)];
            if(Ns > 1) {
                const int jm = j % Ns;
                float2 w1 = tw(jm * (L / (Ns * R)));
                if(INV) w1 = cconj(w1);
                float2 w[R];
                twiddle_powers<R>(w1, w);
#pragma unroll
                for(int k = 1; k < R; ++k) v[q][k] = cmul(v[q][k], w[k]);
            }
            dft<R, INV>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for(int q = 0; q < PER; ++q) {
        const int item = tid + q * NT;
        if(item < C * NB) {
            const int c = item / NB, j = item % NB;
            float2* col = buf + c * padded(L);
            const int jm = j % Ns;
            const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
            for(int k = 0; k < R; ++k) col[pad(idxD + k * Ns)] = v[q][k];
        }
    }
    __syncthreads();
}

// Full FFT of length L (power of two, 16..16384) on C columns in LDS, natural
// order in and out, as radix-16 passes followed by one smaller pass.
template <bool INV, int L, int C, int NT>
__device__ __forceinline__ void fft_lds(float2* buf, const Twiddles& tw)
{
    int Ns = 1;
    constexpr int lg = __builtin_ctz(L);
    constexpr int n16 = lg / 4;
    constexpr int rem = lg % 4;
#pragma unroll
    for(int p = 0; p < n16; ++p) {
        stockham_pass<16, INV, L, C, NT>(buf, Ns, tw);
        Ns *= 16;
    }
    if constexpr(rem == 1) stockham_pass<2, INV, L, C, NT>(buf, Ns, tw);
    if constexpr(rem == 2) stockham_pass<4, INV, L, C, NT>(buf, Ns, tw);
    if constexpr(rem == 3) stockham_pass<8, INV, L, C, NT>(buf, Ns, tw);
}

} // namespace fft
} // namespace sw
