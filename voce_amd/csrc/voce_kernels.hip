// voce_kernels.hip -- CDNA4 (gfx950) kernels of one EVP-FFT increment.
//
// One FFT fixed-point iteration (evolution.c:342-431) is four launches:
//   k_plane_fwd  per (x-plane, stress component): real-to-half-complex FFT
//                along z (two real rows packed per complex row) and complex
//                FFT along y, all in LDS.                       (evolution.c:350-361)
//   k_xgreen     per tile of (y, kz) columns, all 6 components: FFT along
//                x, Green operator applied on the fly from n(k) and C0
//                (init.c:570-615, evolution.c:363-377), inverse FFT along x.
//   k_plane_inv  per (x-plane, component): inverse y FFT, half-complex-to-
//                real z FFT, and the strain update
//                E += dDisGradAvg + IFFT/N (evolution.c:383-394).
//   k_nr         one voxel per lane: the 6x6 Newton-Raphson of UpdateStress
//                (evolution.c:3-168) with StrainRate_eval (constitutive.c:4-72),
//                iterated in the crystal frame of the voxel's TranMat and
//                solved in compliance form by an LDL^T factorisation in
//                registers (in place of LU_inv_66, kinematics.c:282-424);
//                DPP wave sums (partial sums) of the error norms and
//                stress averages per block.  One extra block of the next
//                forward-plane launch (or k_fin_iter after the last launched
//                iteration) sums them in fixed order and runs get_smacro
//                (evolution.c:170-241), the error normalisation and the
//                convergence test (evolution.c:342, 405-428).
// Every launch reads MacroState::active first, so iterations enqueued past
// convergence return at once and the host never blocks inside a loop.
//
// The increment close-out (evolution.c:433-530) is k_xgreen/k_plane_inv in
// skew mode (lattice spin from the summed Green displacements), k_closeout
// (plastic strain, update_orient, harden, diagnostics) and k_fin_incr.
//
// FFT type T is double (FP64 mode) or float (MIXED mode); the constitutive
// update, the reductions and all state stay double.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <math.h>
#include "voce_internal.h"

namespace {

constexpr double RSQ2 = 0.70710678118654744;   // kinematics.c:22-24
constexpr double RSQ3 = 0.57735026918962584;
constexpr double RSQ6 = 0.40824829046386304;
constexpr double TWO_PI_REF = 6.28318530718;    // evp.h:58

// Read-only views in the constant address space: the compiler may then use
// scalar loads for wave-uniform addresses (material constants, C0) even in
// kernels that store to other global memory.  Only for data that no kernel
// writes while the view is in use.
template <typename T> using cptr = const __attribute__((address_space(4))) T *;
template <typename T> __device__ __forceinline__ cptr<T> as_const(const T *p) { return (cptr<T>)p; }
// Global-address-space view of a device pointer: vector loads through it are
// global (not flat) loads, which count against the vector-memory counter
// only, so waiting on them does not also wait on LDS and scalar loads.
template <typename T> using gptr = const __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gptr<T> as_global(const T *p) { return (gptr<T>)p; }
template <typename T> using gmut = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gmut<T> as_global_mut(T *p) { return (gmut<T>)p; }
// The same for the two-component vector types (complex elements, pairs of
// doubles), through the native vector type of the same layout: the HIP
// vector types only bind generic references.
template <typename V2> struct vec2_of;
template <> struct vec2_of<double2> { typedef double t __attribute__((ext_vector_type(2))); };
template <> struct vec2_of<float2> { typedef float t __attribute__((ext_vector_type(2))); };
template <typename V2> __device__ __forceinline__ V2 gload(const V2 *p)
{
    const typename vec2_of<V2>::t v = *(const __attribute__((address_space(1))) typename vec2_of<V2>::t *)p;
    return V2{v.x, v.y};
}
template <typename V2> __device__ __forceinline__ void gstore(V2 *p, V2 a)
{
    typename vec2_of<V2>::t v;
    v.x = a.x;
    v.y = a.y;
    *(__attribute__((address_space(1))) typename vec2_of<V2>::t *)p = v;
}

// Strided element loop e = tid, tid + NT, ... < n whose global loads go out
// U at a time: v = ld(e) for a group of U elements, then st(e, v) for each.
// (In a plain loop every load is waited for before the next is issued, so
// a block spends one memory latency per element it stages.)
template <int NT, int U, typename V, typename LD, typename ST>
__device__ __forceinline__ void grouped(int n, LD ld, ST st)
{
    for (int e0 = (int)threadIdx.x; e0 < n; e0 += U * NT) {
        V v[U];
#pragma unroll
        for (int i = 0; i < U; ++i) {
            const int e = e0 + i * NT;
            if (e < n) v[i] = ld(e);
        }
#pragma unroll
        for (int i = 0; i < U; ++i) {
            const int e = e0 + i * NT;
            if (e < n) st(e, v[i]);
        }
    }
}

template <typename T> struct cplx_of;
template <> struct cplx_of<double> { using t = double2; };
template <> struct cplx_of<float> { using t = float2; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C cmul_mi(C a)
{
    return INV ? C{-a.y, a.x} : C{a.y, -a.x};
}

// ------------------------------------------------------------------
// Radix butterflies (forward sign -1, inverse +1)
// ------------------------------------------------------------------
template <int R, bool INV, typename C> struct Bfly;

template <bool INV, typename C> struct Bfly<2, INV, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV, typename C> struct Bfly<3, INV, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        using T = decltype(v[0].x);
        const T c1 = (T)-0.5;
        const T s1 = INV ? (T)0.86602540378443865 : (T)-0.86602540378443865;
        C a = v[0], b = v[1], c = v[2];
        C t = cadd(b, c);
        C d = csub(b, c);
        v[0] = cadd(a, t);
        C m{a.x + c1 * t.x, a.y + c1 * t.y};
        C r{-s1 * d.y, s1 * d.x};   // i*s1*d
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

template <bool INV, typename C> struct Bfly<4, INV, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        C a0 = cadd(v[0], v[2]);
        C a1 = csub(v[0], v[2]);
        C a2 = cadd(v[1], v[3]);
        C a3 = cmul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    }
};

template <bool INV, typename C> struct Bfly<8, INV, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        using T = decltype(v[0].x);
        const T h = (T)0.70710678118654752;
        // radix-2 on pairs (k, k+4), twiddle, two radix-4 on evens/odds
        C e[4], o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            e[k] = cadd(v[k], v[k + 4]);
            o[k] = csub(v[k], v[k + 4]);
        }
        // o[k] *= W8^k
        {
            C t1 = o[1];
            o[1] = INV ? C{h * (t1.x - t1.y), h * (t1.x + t1.y)} : C{h * (t1.x + t1.y), h * (t1.y - t1.x)};
            o[2] = cmul_mi<INV>(o[2]);
            C t3 = o[3];
            o[3] = INV ? C{-h * (t3.x + t3.y), h * (t3.x - t3.y)} : C{h * (t3.y - t3.x), -h * (t3.x + t3.y)};
        }
        Bfly<4, INV, C>::run(e);
        Bfly<4, INV, C>::run(o);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[2 * k] = e[k];
            v[2 * k + 1] = o[k];
        }
    }
};

template <bool INV, typename C> struct Bfly<16, INV, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        using T = decltype(v[0].x);
        // radix-2 on pairs (k, k+8), twiddle W16^k, two radix-8 on evens/odds
        const T c1 = (T)0.92387953251128674, s1 = (T)0.38268343236508977, h = (T)0.70710678118654752;
        const T wr[8] = {(T)1, c1, h, s1, (T)0, -s1, -h, -c1};
        const T wi[8] = {(T)0, s1, h, c1, (T)1, c1, h, s1};   // W16^k = wr - i wi (forward)
        C e[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            e[k] = cadd(v[k], v[k + 8]);
            o[k] = csub(v[k], v[k + 8]);
        }
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            if (k == 4) {
                o[4] = cmul_mi<INV>(o[4]);
            } else {
                const C w = C{wr[k], INV ? wi[k] : -wi[k]};
                o[k] = cmul(o[k], w);
            }
        }
        Bfly<8, INV, C>::run(e);
        Bfly<8, INV, C>::run(o);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[2 * k] = e[k];
            v[2 * k + 1] = o[k];
        }
    }
};

template <bool INV, typename C> struct Bfly<12, INV, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        using T = decltype(v[0].x);
        // n = n1 + 4 n2, k = 3 k1 + k2: radix-3 over n2 for each n1, twiddle
        // W12^(n1 k2), radix-4 over n1 for each k2
        // (W12^j = c[j] - i s[j] forward, c[j] = cos(pi j / 6))
        const T c[7] = {(T)1, (T)0.86602540378443865, (T)0.5, (T)0, (T)-0.5, (T)-0.86602540378443865, (T)-1};
        const T s[7] = {(T)0, (T)0.5, (T)0.86602540378443865, (T)1, (T)0.86602540378443865, (T)0.5, (T)0};
        C y[3][4];
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) {
            C t[3] = {v[n1], v[n1 + 4], v[n1 + 8]};
            Bfly<3, INV, C>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) y[k2][n1] = t[k2];
        }
#pragma unroll
        for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
            for (int k2 = 1; k2 < 3; ++k2) {
                const int j = n1 * k2;
                if (j == 3) y[k2][n1] = cmul_mi<INV>(y[k2][n1]);
                else if (j == 6) y[k2][n1] = C{-y[k2][n1].x, -y[k2][n1].y};
                else y[k2][n1] = cmul(y[k2][n1], C{c[j], INV ? s[j] : -s[j]});
            }
#pragma unroll
        for (int k2 = 0; k2 < 3; ++k2) {
            Bfly<4, INV, C>::run(y[k2]);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) v[3 * k1 + k2] = y[k2][k1];
        }
    }
};

// Odd-prime butterflies (R = 5, 7) in the symmetric form: with
// t_j = v_j + v_(R-j), u_j = v_j - v_(R-j) (j = 1..(R-1)/2),
// X_0 = v_0 + sum t_j, and for k = 1..(R-1)/2
// X_k = m_k -/+ i n_k, X_(R-k) = m_k +/- i n_k (forward/inverse) with
// m_k = v_0 + sum_j cos(2 pi j k / R) t_j, n_k = sum_j sin(2 pi j k / R) u_j.
template <int R> struct OddTrig;
template <> struct OddTrig<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494742, -0.80901699437494742, -0.80901699437494742,
                                    0.30901699437494742};
    static constexpr double s[5] = {0.0, 0.95105651629515357, 0.58778525229247313, -0.58778525229247313,
                                    -0.95105651629515357};
};
template <> struct OddTrig<7> {
    static constexpr double c[7] = {1.0, 0.62348980185873353, -0.22252093395631440, -0.90096886790241913,
                                    -0.90096886790241913, -0.22252093395631440, 0.62348980185873353};
    static constexpr double s[7] = {0.0, 0.78183148246802981, 0.97492791218182361, 0.43388373911755812,
                                    -0.43388373911755812, -0.97492791218182361, -0.78183148246802981};
};

template <int R, bool INV, typename C>
__device__ __forceinline__ void bfly_odd(C *v)
{
    using T = decltype(v[0].x);
    constexpr int H = (R - 1) / 2;
    C t[H + 1], u[H + 1];
    C x0 = v[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        t[j] = cadd(v[j], v[R - j]);
        u[j] = csub(v[j], v[R - j]);
        x0 = cadd(x0, t[j]);
    }
    C out[R];
    out[0] = x0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        C m = v[0], n = C{(T)0, (T)0};
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const T cj = (T)OddTrig<R>::c[(j * k) % R], sj = (T)OddTrig<R>::s[(j * k) % R];
            m = C{m.x + cj * t[j].x, m.y + cj * t[j].y};
            n = C{n.x + sj * u[j].x, n.y + sj * u[j].y};
        }
        if (!INV) {
            out[k] = C{m.x + n.y, m.y - n.x};
            out[R - k] = C{m.x - n.y, m.y + n.x};
        } else {
            out[k] = C{m.x - n.y, m.y + n.x};
            out[R - k] = C{m.x + n.y, m.y - n.x};
        }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
}
template <bool INV, typename C> struct Bfly<5, INV, C> {
    __device__ __forceinline__ static void run(C *v) { bfly_odd<5, INV, C>(v); }
};
template <bool INV, typename C> struct Bfly<7, INV, C> {
    __device__ __forceinline__ static void run(C *v) { bfly_odd<7, INV, C>(v); }
};

// Index arithmetic of the FFT kernels.  Every index there is below 2^22,
// so products use the full-rate 24-bit multiply, and a / d comes from a
// float reciprocal with one correction step ((float)a is exact and a * (1/d)
// is within 1/2 of a/d, so the truncated quotient is off by at most one)
// instead of the integer division sequence with its quarter-rate multiplies.
__device__ __forceinline__ int mul24(int a, int b) { return (int)__umul24((unsigned)a, (unsigned)b); }

struct FastDiv {
    int d;
    float r;
    __device__ __forceinline__ explicit FastDiv(int d_) : d(d_), r(1.0f / (float)d_) {}
    // a / d, with a % d in rem (0 <= a < 2^22)
    __device__ __forceinline__ int div(int a, int &rem) const
    {
        int q = (int)((float)a * r);
        int m = a - mul24(q, d);
        if (m < 0) {
            --q;
            m += d;
        } else if (m >= d) {
            ++q;
            m -= d;
        }
        rem = m;
        return q;
    }
};

// ------------------------------------------------------------------
// One Stockham stage over `nlines` lines of length N held in LDS.
// Element q of line l lives at s[l*lstride + q*estride].  All loads of a
// stage complete (barrier) before any store, so the stage runs in place.
// LF: consecutive lanes take consecutive lines: with an odd line stride
// (the x pass) a stage's loads and stores hit distinct banks.
// ------------------------------------------------------------------
template <typename C, int R, int NT, int MAXEL, bool INV, bool LF>
__device__ __forceinline__ void fft_stage(C *s, int nlines, int N, int Ns, int lstride, int estride,
                                          const C *tw)
{
    constexpr int BPT = (MAXEL / R + NT - 1) / NT;
    const int nbl = N / R;
    const int nb = nlines * nbl;
    // all index products stay below 2^24: 24-bit multiplies
    const FastDiv split(LF ? nlines : nbl), byns(Ns);
    C v[BPT][R];
    int base[BPT], jj[BPT];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = (int)threadIdx.x + i * NT;
        if (b < nb) {
            int l, j, m;
            if (LF) { j = split.div(b, m); l = m; }
            else { l = split.div(b, m); j = m; }
            base[i] = mul24(l, lstride);
            jj[i] = j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = s[base[i] + mul24(j + r * nbl, estride)];
        }
    }
    __syncthreads();
    const int twstep = N / (Ns * R);
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = (int)threadIdx.x + i * NT;
        if (b < nb) {
            const int j = jj[i];
            int k;
            const int jd = byns.div(j, k);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[mul24(r * k, twstep)];
                if (INV) w.y = -w.y;
                v[i][r] = cmul(v[i][r], w);
            }
            Bfly<R, INV, C>::run(v[i]);
            const int d = mul24(jd, Ns * R) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) s[base[i] + mul24(d + r * Ns, estride)] = v[i][r];
        }
    }
    __syncthreads();
}

// Mixed-radix plan: radix-3, -5 and -7 stages, then radix 8/4/2 for the
// power of two (r8 = 0: radix 4/2 only; the choice is tuned per grid size).
template <typename C, int NT, int MAXEL, bool INV, bool LF>
__device__ void fft_lines(C *s, int nlines, int N, int lstride, int estride, const C *tw, int r8)
{
    int Ns = 1, rem = N;
    while (rem % 3 == 0) {
        fft_stage<C, 3, NT, MAXEL, INV, LF>(s, nlines, N, Ns, lstride, estride, tw);
        Ns *= 3;
        rem /= 3;
    }
    while (rem % 5 == 0) {
        fft_stage<C, 5, NT, MAXEL, INV, LF>(s, nlines, N, Ns, lstride, estride, tw);
        Ns *= 5;
        rem /= 5;
    }
    while (rem % 7 == 0) {
        fft_stage<C, 7, NT, MAXEL, INV, LF>(s, nlines, N, Ns, lstride, estride, tw);
        Ns *= 7;
        rem /= 7;
    }
    int a = 0;
    while ((1 << a) < rem) ++a;
    int n8 = a / 3, n4 = 0, n2 = 0;
    if (a % 3 == 2) n4 = 1;
    if (a % 3 == 1) {
        if (n8 >= 1) { n8 -= 1; n4 = 2; }
        else n2 = 1;
    }
    if (!r8) {
        n8 = 0;
        n4 = a / 2;
        n2 = a % 2;
    }
    for (int i = 0; i < n8; ++i) { fft_stage<C, 8, NT, MAXEL, INV, LF>(s, nlines, N, Ns, lstride, estride, tw); Ns *= 8; }
    for (int i = 0; i < n4; ++i) { fft_stage<C, 4, NT, MAXEL, INV, LF>(s, nlines, N, Ns, lstride, estride, tw); Ns *= 4; }
    for (int i = 0; i < n2; ++i) { fft_stage<C, 2, NT, MAXEL, INV, LF>(s, nlines, N, Ns, lstride, estride, tw); Ns *= 2; }
}

// Kernels take a table of RVE descriptors and an optional block map.  A
// single-RVE launch passes a one-entry table and no map (local block =
// blockIdx.x); a batched launch maps every block to (RVE, local block), so
// one grid covers RVEs of different sizes.
#define RVE_PROLOGUE(LB)                                                     \
    int rid_ = 0, LB = (int)blockIdx.x;                                      \
    if (bmap) {                                                              \
        const int2 m_ = bmap[blockIdx.x];                                    \
        rid_ = m_.x;                                                         \
        LB = m_.y;                                                           \
    }                                                                        \
    const RveDev &r = rtab[rid_]

template <int NT> __device__ __forceinline__ void finalize_block(const RveDev &r);

// Local block index lb of a forward plane launch: its finaliser block
// (lb == 6 nx) is the launch's first block and the plane blocks follow, so
// the finaliser's row sums overlap them; dispatched last, they added their
// whole duration to the launch (22 us of row sums at 128^3).
__device__ __forceinline__ int fin_first(int lb0, int nx) { return lb0 == 0 ? 6 * nx : lb0 - 1; }

// The macro state is written only by single-thread finalisers in other
// launches (or, racing harmlessly, by the finaliser block of a forward plane
// launch, whose other blocks only write scratch spectra), so kernels read it
// through the scalar cache: a vector load would be waited for together with
// every load issued before it.
__device__ __forceinline__ bool iter_active(const MacroState *ms)
{
    return as_const(ms)->active != 0;
}

// The increment's spin is needed by update_orient from step 2
// (evolution.c:500), by the creep predictor and by the diagnostics.
__device__ __forceinline__ bool rve_need_skew(const MacroState *ms)
{
    const cptr<MacroState> m = as_const(ms);
    return (m->update_flag == 1 && m->istep > 1) || m->creep || m->diag;
}

// ------------------------------------------------------------------
// Forward plane kernel: S6 component c of plane x -> SPEC[c][x][:][:]
// grid (nx, 6), dynamic LDS (ny*nzh + ny + nz) complex
// ------------------------------------------------------------------
template <typename T, int NT, int MAXEL>
__global__ __launch_bounds__(NT) void k_plane_fwd(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    using C = typename cplx_of<T>::t;
    RVE_PROLOGUE(lb0);
    const int lb = fin_first(lb0, r.nx);
    if (lb == 6 * r.nx) {   // the previous iteration's finaliser
        finalize_block<NT>(r);
        return;
    }
    // a finaliser may be deciding convergence beside this block: either
    // value of the flag is fine here, the output is scratch
    if (!iter_active(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx, ny = r.ny, nz = r.nz, nzh = r.nzh;
    const int RS = 2 * nzh;   // packed z rows: two real rows as one complex row
    // an odd last row is packed with a row of zeros
    const int np = (ny + 1) >> 1;
    C *twy = s + 2 * np * nzh;
    C *twz = twy + ny;
    const int c = lb / nx, x = lb - c * nx, tid = threadIdx.x;
    const C *gty = (const C *)r.tw_y;
    const C *gtz = (const C *)r.tw_z;
    for (int i = tid; i < ny; i += NT) twy[i] = gty[i];
    for (int i = tid; i < nz; i += NT) twz[i] = gtz[i];
    const gptr<double> src = as_global(r.S6 + (long)c * r.V + (long)x * ny * nz);
    T *sr = (T *)s;
    const FastDiv bynz(nz), bynzh(nzh);
    grouped<NT, 16, double>(ny * nz, [&](int e) { return src[e]; }, [&](int e, double v) {
        int z;
        const int y = bynz.div(e, z);
        sr[2 * ((y >> 1) * RS + z) + (y & 1)] = (T)v;
    });
    if (ny & 1)
        for (int z = tid; z < nz; z += NT) sr[2 * ((np - 1) * RS + z) + 1] = (T)0;
    __syncthreads();
    fft_lines<C, NT, MAXEL, false, false>(s, np, nz, RS, 1, twz, r.pl_r8);
    // split packed rows a + i b into the half spectra of a and b
    {
        constexpr int SPT = (MAXEL / 2 + NT - 1) / NT;
        C A[SPT], B[SPT];
        const int nit = np * nzh;
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int it = tid + i * NT;
            if (it < nit) {
                int k;
                const int p = bynzh.div(it, k);
                const C z0 = s[p * RS + k];
                const int km = (k == 0) ? 0 : nz - k;
                const C z1 = s[p * RS + km];
                A[i] = C{(T)0.5 * (z0.x + z1.x), (T)0.5 * (z0.y - z1.y)};
                B[i] = C{(T)0.5 * (z0.y + z1.y), (T)0.5 * (z1.x - z0.x)};
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int it = tid + i * NT;
            if (it < nit) {
                int k;
                const int p = bynzh.div(it, k);
                s[(2 * p) * nzh + k] = A[i];
                if (2 * p + 1 < ny) s[(2 * p + 1) * nzh + k] = B[i];
            }
        }
        __syncthreads();
    }
    fft_lines<C, NT, MAXEL, false, true>(s, nzh, ny, 1, nzh, twy, r.pl_r8);
    // column-blocked half spectrum: runs of TB columns (128 bytes) per x
    constexpr int TB = 128 / (int)sizeof(C);
    C *dst = (C *)r.spec + (long)c * r.nspec;
    for (int e = tid; e < ny * nzh; e += NT) dst[((long)(e / TB) * nx + x) * TB + (e % TB)] = s[e];
}

// ------------------------------------------------------------------
// Inverse plane kernel.  mode 0: E6[c] += shift[c] + result (or, in the
// first iteration of a creep increment, E6[c] = creep[c] + result);
// mode 1: WF[c] = result (skew recovery).  grid (nx, ncomp)
// ------------------------------------------------------------------
template <typename T, int NT, int MAXEL>
__global__ __launch_bounds__(NT) void k_plane_inv(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap, int mode)
{
    using C = typename cplx_of<T>::t;
    RVE_PROLOGUE(lb);
    if (mode == 0 ? !iter_active(r.ms) : !rve_need_skew(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx, ny = r.ny, nz = r.nz, nzh = r.nzh;
    const int RS = 2 * nzh;   // packed z rows, as in k_plane_fwd
    const int np = (ny + 1) >> 1;
    C *twy = s + 2 * np * nzh;
    C *twz = twy + ny;
    const int c = lb / nx, x = lb - c * nx, tid = threadIdx.x;
    const FastDiv bynz(nz);
    const C *gty = (const C *)r.tw_y;
    const C *gtz = (const C *)r.tw_z;
    for (int i = tid; i < ny; i += NT) twy[i] = gty[i];
    for (int i = tid; i < nz; i += NT) twz[i] = gtz[i];
    constexpr int TB = 128 / (int)sizeof(C);
    const C *srcp = (const C *)r.spec + (long)c * r.nspec;
    grouped<NT, 8, C>(ny * nzh, [&](int e) { return srcp[((long)(e / TB) * nx + x) * TB + (e % TB)]; },
                      [&](int e, C v) { s[e] = v; });
    __syncthreads();
    fft_lines<C, NT, MAXEL, true, true>(s, nzh, ny, 1, nzh, twy, r.pl_r8);
    // pack the half spectra of rows 2p (a) and 2p+1 (b) into a full complex
    // row Z = A + iB, A and B Hermitian-extended
    {
        constexpr int SPT = (MAXEL + NT - 1) / NT;
        C Z[SPT];
        const int nit = np * nz;
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int it = tid + i * NT;
            if (it < nit) {
                int k;
                const int p = bynz.div(it, k);
                C A, B;
                const bool hb = 2 * p + 1 < ny;   // an odd last row pairs with zeros
                if (k <= nz / 2) {
                    A = s[(2 * p) * nzh + k];
                    B = hb ? s[(2 * p + 1) * nzh + k] : C{(T)0, (T)0};
                    if (k == 0 || 2 * k == nz) { A.y = (T)0; B.y = (T)0; }
                } else {
                    A = s[(2 * p) * nzh + nz - k];
                    B = hb ? s[(2 * p + 1) * nzh + nz - k] : C{(T)0, (T)0};
                    A.y = -A.y;
                    B.y = -B.y;
                }
                Z[i] = C{A.x - B.y, A.y + B.x};
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int it = tid + i * NT;
            if (it < nit) {
                int k;
                const int p = bynz.div(it, k);
                s[p * RS + k] = Z[i];
            }
        }
        __syncthreads();
    }
    fft_lines<C, NT, MAXEL, true, false>(s, np, nz, RS, 1, twz, r.pl_r8);
    const T *sr = (const T *)s;
    const long off = (long)c * r.V + (long)x * ny * nz;
    if (mode == 0) {
        const cptr<MacroState> ms = as_const(r.ms);
        const gmut<double> E = as_global_mut(r.E6 + off);
        const bool creep_set = (ms->creep != 0) && (ms->iter == 0);
        const double sh = creep_set ? 0.0 : ms->shift6[c];
        const double cr = ms->creep_sym6[c];
        grouped<NT, 16, double>(ny * nz, [&](int e) { return creep_set ? 0.0 : E[e]; }, [&](int e, double ev) {
            int z;
            const int y = bynz.div(e, z);
            const double val = (double)sr[2 * ((y >> 1) * RS + z) + (y & 1)];
            E[e] = creep_set ? cr + (sh + val) : ev + (sh + val);
        });
        // diagnostics: the running sum of the Green updates (the reference's
        // fluctuation, evolution.c:392) for the components new_position reads
        const int fs = (c == 1) ? 0 : (c == 2) ? 1 : (c == 3) ? 2 : (c == 5) ? 3 : -1;
        if (r.FLU && fs >= 0) {
            double *F = r.FLU + (long)fs * r.V + (long)x * ny * nz;
            for (int e = tid; e < ny * nz; e += NT) {
                int z;
                const int y = bynz.div(e, z);
                F[e] += (double)sr[2 * ((y >> 1) * RS + z) + (y & 1)];
            }
        }
    } else {
        double *W = r.WF + off;
        double *F = (r.FLU && c < 2) ? r.FLU + (long)(4 + c) * r.V + (long)x * ny * nz : nullptr;
        for (int e = tid; e < ny * nz; e += NT) {
            int z;
            const int y = bynz.div(e, z);
            const double val = (double)sr[2 * ((y >> 1) * RS + z) + (y & 1)];
            W[e] = val;
            if (F) F[e] += val;
        }
    }
}

// ------------------------------------------------------------------
// Split plane passes (plane variant 5), for planes too large for one
// block's LDS (any Ny, Nz up to 512).  The z and y transforms of a plane are
// two launches; a block owns one chunk of lines (that fit MAXEL_SP elements
// of LDS) of one (component, x-plane): blockIdx.x maps to the (component,
// x-plane) as in the whole-plane kernels (so the block maps are the same) and
// blockIdx.y is the chunk.  Between the two launches the half
// spectrum of z row y sits in the spectrum array at column y * nzh + kz (the
// column-blocked layout of the final spectrum, with y in place of ky); the
// y pass transforms each kz column in place.  An odd Ny pairs its last row
// with a row of zeros, as in the whole-plane kernels.
// ------------------------------------------------------------------
// 2048 elements per chunk (32 KB of FP64): four blocks per CU; 4096 (two
// blocks) and 1024 (more chunks per block) measured slower at 256^3
constexpr int NT_SP = 256, MAXEL_SP = 2048;
constexpr int SP_MAXN = 512;   // largest Ny, Nz of the split passes

template <typename C>
__device__ __forceinline__ long spec_at(int col, int x, int nx)
{
    constexpr int TB = 128 / (int)sizeof(C);
    return ((long)(col / TB) * nx + x) * TB + (col % TB);
}

// forward z pass: S6 rows (packed in pairs) -> half spectra; the extra block
// lb == 6 nx finalises the previous iteration
template <typename T>
__global__ __launch_bounds__(NT_SP) void k_zfwd_s(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    using C = typename cplx_of<T>::t;
    RVE_PROLOGUE(lb0);
    const int lb = fin_first(lb0, r.nx);
    if (lb == 6 * r.nx) {
        if (blockIdx.y == 0) finalize_block<NT_SP>(r);
        return;
    }
    if (!iter_active(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx, ny = r.ny, nz = r.nz, nzh = r.nzh, tid = threadIdx.x;
    const int RS = 2 * nzh, np = (ny + 1) >> 1;
    C *twz = s + MAXEL_SP;
    for (int i = tid; i < nz; i += NT_SP) twz[i] = gload((const C *)r.tw_z + i);
    const int c = lb / nx, x = lb - c * nx;
    const gptr<double> src = as_global(r.S6 + (long)c * r.V + (long)x * ny * nz);
    C *dst = (C *)r.spec + (long)c * r.nspec;
    T *sr = (T *)s;
    const FastDiv bynz(nz), bynzh(nzh);
    const int ppc = MAXEL_SP / RS;   // row pairs per chunk
    {
        const int p0 = (int)blockIdx.y * ppc;
        if (p0 >= np) return;
        const int npc = min(ppc, np - p0);
        const int nrow = min(2 * npc, ny - 2 * p0);
        grouped<NT_SP, 16, double>(nrow * nz, [&](int e) { return src[(long)2 * p0 * nz + e]; },
                                   [&](int e, double v) {
                                       int z;
                                       const int y = bynz.div(e, z);
                                       sr[2 * ((y >> 1) * RS + z) + (y & 1)] = (T)v;
                                   });
        if (nrow < 2 * npc)
            for (int z = tid; z < nz; z += NT_SP) sr[2 * ((npc - 1) * RS + z) + 1] = (T)0;
        __syncthreads();
        fft_lines<C, NT_SP, MAXEL_SP, false, false>(s, npc, nz, RS, 1, twz, r.pl_r8);
        // split packed rows a + i b into the half spectra of a and b, out to
        // columns y nzh + k
        for (int it = tid; it < npc * nzh; it += NT_SP) {
            int k;
            const int p = bynzh.div(it, k);
            const C z0 = s[p * RS + k];
            const C z1 = s[p * RS + ((k == 0) ? 0 : nz - k)];
            const int y = 2 * (p0 + p);
            gstore(dst + spec_at<C>(y * nzh + k, x, nx), C{(T)0.5 * (z0.x + z1.x), (T)0.5 * (z0.y - z1.y)});
            if (y + 1 < ny)
                gstore(dst + spec_at<C>((y + 1) * nzh + k, x, nx), C{(T)0.5 * (z0.y + z1.y), (T)0.5 * (z1.x - z0.x)});
        }
        __syncthreads();
    }
}

// y pass, forward or inverse, in place on the kz columns of one plane
template <typename T, bool INV>
__global__ __launch_bounds__(NT_SP) void k_y_s(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap, int mode)
{
    using C = typename cplx_of<T>::t;
    RVE_PROLOGUE(lb);
    if (mode == 0 ? !iter_active(r.ms) : !rve_need_skew(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx, ny = r.ny, nzh = r.nzh, tid = threadIdx.x;
    C *twy = s + MAXEL_SP;
    for (int i = tid; i < ny; i += NT_SP) twy[i] = gload((const C *)r.tw_y + i);
    const int c = lb / nx, x = lb - c * nx;
    C *sp = (C *)r.spec + (long)c * r.nspec;
    const int kpc = MAXEL_SP / ny;   // kz columns per chunk
    {
        const int k0 = (int)blockIdx.y * kpc;
        if (k0 >= nzh) return;
        const int nk = min(kpc, nzh - k0);
        const FastDiv bynk(nk);
        // element e = y nk + j: consecutive lanes on consecutive columns
        grouped<NT_SP, 8, C>(ny * nk, [&](int e) {
            int j;
            const int y = bynk.div(e, j);
            return gload(sp + spec_at<C>(y * nzh + k0 + j, x, nx));
        }, [&](int e, C v) {
            int j;
            const int y = bynk.div(e, j);
            s[j * ny + y] = v;
        });
        __syncthreads();
        fft_lines<C, NT_SP, MAXEL_SP, INV, false>(s, nk, ny, ny, 1, twy, r.pl_r8);
        for (int e = tid; e < ny * nk; e += NT_SP) {
            int j;
            const int y = bynk.div(e, j);
            gstore(sp + spec_at<C>(y * nzh + k0 + j, x, nx), s[j * ny + y]);
        }
        __syncthreads();
    }
}

// inverse z pass: half spectra of row pairs -> real rows; mode 0: E6[c] +=
// shift[c] + result (creep: E6[c] = creep[c] + result); mode 1: WF[c] =
// result
template <typename T>
__global__ __launch_bounds__(NT_SP) void k_zinv_s(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap, int mode)
{
    using C = typename cplx_of<T>::t;
    RVE_PROLOGUE(lb);
    if (mode == 0 ? !iter_active(r.ms) : !rve_need_skew(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx, ny = r.ny, nz = r.nz, nzh = r.nzh, tid = threadIdx.x;
    const int RS = 2 * nzh, np = (ny + 1) >> 1;
    C *twz = s + MAXEL_SP;
    for (int i = tid; i < nz; i += NT_SP) twz[i] = gload((const C *)r.tw_z + i);
    const int c = lb / nx, x = lb - c * nx;
    const C *sp = (const C *)r.spec + (long)c * r.nspec;
    const FastDiv bynz(nz), bynzh(nzh);
    const cptr<MacroState> ms = as_const(r.ms);
    const bool creep_set = mode == 0 && (ms->creep != 0) && (ms->iter == 0);
    const double sh = (mode == 0 && !creep_set) ? ms->shift6[c] : 0.0;
    const double cr = mode == 0 ? ms->creep_sym6[c] : 0.0;
    const long off = (long)c * r.V + (long)x * ny * nz;
    const int fs = (c == 1) ? 0 : (c == 2) ? 1 : (c == 3) ? 2 : (c == 5) ? 3 : -1;
    double *F = nullptr;
    if (r.FLU && mode == 0 && fs >= 0) F = r.FLU + (long)fs * r.V + (long)x * ny * nz;
    if (r.FLU && mode == 1 && c < 2) F = r.FLU + (long)(4 + c) * r.V + (long)x * ny * nz;
    const T *sr = (const T *)s;
    const int ppc = MAXEL_SP / RS;
    {
        const int p0 = (int)blockIdx.y * ppc;
        if (p0 >= np) return;
        const int npc = min(ppc, np - p0);
        const int nrow = min(2 * npc, ny - 2 * p0);
        // the chunk's half-spectrum rows, row y - 2 p0 at (y - 2 p0) nzh
        grouped<NT_SP, 8, C>(nrow * nzh, [&](int e) {
            int k;
            const int yl = bynzh.div(e, k);
            return gload(sp + spec_at<C>((2 * p0 + yl) * nzh + k, x, nx));
        }, [&](int e, C v) { s[e] = v; });
        __syncthreads();
        // pack rows 2p (a) and 2p+1 (b) into Z = A + iB, A and B
        // Hermitian-extended (an odd last row pairs with zeros)
        constexpr int SPT = (MAXEL_SP + NT_SP - 1) / NT_SP;
        C Z[SPT];
        const int nit = npc * nz;
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int it = tid + i * NT_SP;
            if (it < nit) {
                int k;
                const int p = bynz.div(it, k);
                const bool hb = 2 * p + 1 < nrow;
                C A, B;
                if (k <= nz / 2) {
                    A = s[(2 * p) * nzh + k];
                    B = hb ? s[(2 * p + 1) * nzh + k] : C{(T)0, (T)0};
                    if (k == 0 || 2 * k == nz) { A.y = (T)0; B.y = (T)0; }
                } else {
                    A = s[(2 * p) * nzh + nz - k];
                    B = hb ? s[(2 * p + 1) * nzh + nz - k] : C{(T)0, (T)0};
                    A.y = -A.y;
                    B.y = -B.y;
                }
                Z[i] = C{A.x - B.y, A.y + B.x};
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int it = tid + i * NT_SP;
            if (it < nit) {
                int k;
                const int p = bynz.div(it, k);
                s[p * RS + k] = Z[i];
            }
        }
        __syncthreads();
        fft_lines<C, NT_SP, MAXEL_SP, true, false>(s, npc, nz, RS, 1, twz, r.pl_r8);
        // rows 2 p0 .. 2 p0 + nrow - 1 of the plane
        const long o = (long)2 * p0 * nz;
        if (mode == 0) {
            const gmut<double> E = as_global_mut(r.E6 + off + o);
            grouped<NT_SP, 16, double>(nrow * nz, [&](int e) { return creep_set ? 0.0 : E[e]; }, [&](int e, double ev) {
                int z;
                const int y = bynz.div(e, z);
                const double val = (double)sr[2 * ((y >> 1) * RS + z) + (y & 1)];
                E[e] = creep_set ? cr + (sh + val) : ev + (sh + val);
            });
        } else {
            const gmut<double> W = as_global_mut(r.WF + off + o);
            for (int e = tid; e < nrow * nz; e += NT_SP) {
                int z;
                const int y = bynz.div(e, z);
                W[e] = (double)sr[2 * ((y >> 1) * RS + z) + (y & 1)];
            }
        }
        if (F) {
            // diagnostics: the running sum of the Green updates (evolution.c:392)
            for (int e = tid; e < nrow * nz; e += NT_SP) {
                int z;
                const int y = bynz.div(e, z);
                F[o + e] += (double)sr[2 * ((y >> 1) * RS + z) + (y & 1)];
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------
// Line-parallel plane kernels for square planes (Ny = Nz = N, N = 16, 32,
// 48, 64, 128): every line of N points (a packed z-row pair, then a y
// column of the half spectrum) belongs to a group of L lanes of one wave,
// R = N / L points per lane.  A line is transformed by an R-point DFT in
// registers, the twiddles W_N^(n1 k1), one exchange through the line's own
// LDS slots and L-point DFTs in registers.  The exchange only involves the
// lanes of one wave (whose LDS operations are performed in order), so a
// plane pass has one block barrier, between its z and y halves, and no
// index arithmetic beyond constant strides.  Used when the tuner picks
// plane variant 4 for a grid.
// ------------------------------------------------------------------
template <int N> struct P2Plan;
template <> struct P2Plan<16> { static constexpr int L = 4, R = 4, NT = 64; };
template <> struct P2Plan<32> { static constexpr int L = 4, R = 8, NT = 128; };
template <> struct P2Plan<48> { static constexpr int L = 4, R = 12, NT = 128; };
template <> struct P2Plan<64> { static constexpr int L = 8, R = 8, NT = 320; };
template <> struct P2Plan<128> { static constexpr int L = 8, R = 16, NT = 576; };

// Orders the LDS accesses of one wave (no instruction is emitted: the
// accesses of a wave are performed in program order; this keeps the
// compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// N-point DFT (forward, or inverse with INV) of the line held by the L lanes
// of a group: on entry the lane with index n1 in its group holds x[n1 + L r]
// in v[r]; on exit it holds X[n1 + L j + R k2] in v[j L + k2] (j < R / L,
// k2 < L).  line(i) is the address of the i-th of the line's N LDS slots
// (owned by the group; every lane must have read what it needs of them
// before the call).  tw[k1] = W_N^(n1 k1), forward sign.
//   X[k1 + R k2] = sum_n1 W_L^(n1 k2) W_N^(n1 k1) sum_r x[n1 + L r] W_R^(r k1)
template <typename C, int N, bool INV, typename LINE>
__device__ __forceinline__ void line_dft(C (&v)[P2Plan<N>::R], const C (&tw)[P2Plan<N>::R], int n1, LINE line)
{
    constexpr int L = P2Plan<N>::L, R = P2Plan<N>::R;
    Bfly<R, INV, C>::run(v);
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) {
        C w = tw[k1];
        if (INV) w.y = -w.y;
        v[k1] = cmul(v[k1], w);
    }
    wave_lds_sync();
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) *line(k1 * L + n1) = v[k1];
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < R / L; ++j)
#pragma unroll
        for (int n = 0; n < L; ++n) v[j * L + n] = *line((n1 + L * j) * L + n);
#pragma unroll
    for (int j = 0; j < R / L; ++j) Bfly<L, INV, C>::run(&v[j * L]);
}

// index of the line element a lane holds in v[o] after line_dft
template <int N> __device__ __forceinline__ constexpr int line_out_index(int n1, int o)
{
    return n1 + P2Plan<N>::L * (o / P2Plan<N>::L) + P2Plan<N>::R * (o % P2Plan<N>::L);
}

// Forward plane pass (the work of k_plane_fwd) for Ny = Nz = N.
template <typename T, int N>
__global__ __launch_bounds__(P2Plan<N>::NT) void k_plane_fwd_p2(const RveDev *__restrict__ rtab,
                                                                const int2 *__restrict__ bmap)
{
    using C = typename cplx_of<T>::t;
    constexpr int L = P2Plan<N>::L, R = P2Plan<N>::R, NT = P2Plan<N>::NT;
    constexpr int NZH = N / 2 + 1, RS = 2 * NZH, KT = N / (2 * L);
    RVE_PROLOGUE(lb0);
    const int lb = fin_first(lb0, r.nx);
    if (lb == 6 * r.nx) {   // the previous iteration's finaliser
        finalize_block<NT>(r);
        return;
    }
    if (!iter_active(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;   // the plane: row y at y NZH (rows 2p, 2p+1 = packed row p at p RS)
    const int nx = r.nx;
    const int c = lb / nx, x = lb - c * nx;
    const int grp = (int)threadIdx.x / L, n1 = (int)threadIdx.x % L;
    C tw[R];
    {
        const C *gtw = (const C *)r.tw_z;   // tw_y holds the same table (Ny = Nz)
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) tw[k1] = gload(gtw + n1 * k1);
    }
    // z: packed row pair p = grp, two real rows as one complex row
    if (grp < N / 2) {
        // the two rows in pairs of points, consecutive lanes on consecutive
        // pairs (whole cache lines per instruction), into the line's slots in
        // natural order, then read back in the layout line_dft takes
        const double2 *src = (const double2 *)(r.S6 + (long)c * r.V + ((long)x * N + 2 * grp) * N);
        C *line = s + grp * RS;
        double2 pa[R / 2], pb[R / 2];
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
            pa[q] = gload(src + n1 + L * q);
            pb[q] = gload(src + N / 2 + n1 + L * q);
        }
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
            const int z = 2 * (n1 + L * q);
            line[z] = C{(T)pa[q].x, (T)pb[q].x};
            line[z + 1] = C{(T)pa[q].y, (T)pb[q].y};
        }
        wave_lds_sync();
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = line[n1 + L * q];
        line_dft<C, N, false>(v, tw, n1, [&](int i) { return line + i; });
        // split Z = a + i b into the half spectra of a and b:
        // A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i, k <= N/2
        wave_lds_sync();
#pragma unroll
        for (int o = 0; o < R; ++o) line[line_out_index<N>(n1, o)] = v[o];
        wave_lds_sync();
        C A[KT + 1], B[KT + 1];
#pragma unroll
        for (int t = 0; t <= KT; ++t) {
            const int k = n1 + L * t;
            if (k <= N / 2) {
                const C z0 = line[k], z1 = line[(N - k) % N];
                A[t] = C{(T)0.5 * (z0.x + z1.x), (T)0.5 * (z0.y - z1.y)};
                B[t] = C{(T)0.5 * (z0.y + z1.y), (T)0.5 * (z1.x - z0.x)};
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t <= KT; ++t) {
            const int k = n1 + L * t;
            if (k <= N / 2) {
                line[k] = A[t];
                line[NZH + k] = B[t];
            }
        }
    }
    __syncthreads();
    // y: half-spectrum column kz = grp
    if (grp < NZH) {
        C *col = s + grp;
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = col[(n1 + L * q) * NZH];
        line_dft<C, N, false>(v, tw, n1, [&](int i) { return col + i * NZH; });
        constexpr int TB = 128 / (int)sizeof(C);
        C *dst = (C *)r.spec + (long)c * r.nspec;
#pragma unroll
        for (int o = 0; o < R; ++o) {
            const int e = line_out_index<N>(n1, o) * NZH + grp;
            gstore(dst + ((long)(e / TB) * nx + x) * TB + (e % TB), v[o]);
        }
    }
}

// Inverse plane pass (the work of k_plane_inv, both modes) for Ny = Nz = N.
template <typename T, int N>
__global__ __launch_bounds__(P2Plan<N>::NT) void k_plane_inv_p2(const RveDev *__restrict__ rtab,
                                                                const int2 *__restrict__ bmap, int mode)
{
    using C = typename cplx_of<T>::t;
    constexpr int L = P2Plan<N>::L, R = P2Plan<N>::R;
    constexpr int NZH = N / 2 + 1, RS = 2 * NZH;
    RVE_PROLOGUE(lb);
    if (mode == 0 ? !iter_active(r.ms) : !rve_need_skew(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx;
    const int c = lb / nx, x = lb - c * nx;
    const int grp = (int)threadIdx.x / L, n1 = (int)threadIdx.x % L;
    const cptr<MacroState> ms = as_const(r.ms);
    const bool creep_set = mode == 0 && (ms->creep != 0) && (ms->iter == 0);
    const long off = (long)c * r.V + ((long)x * N + 2 * grp) * N;   // row 2p of this group
    C tw[R];
    {
        const C *gtw = (const C *)r.tw_z;
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) tw[k1] = gload(gtw + n1 * k1);
    }
    // y: half-spectrum column kz = grp, inverse DFT along y
    if (grp < NZH) {
        constexpr int TB = 128 / (int)sizeof(C);
        const C *srcp = (const C *)r.spec + (long)c * r.nspec;
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int e = (n1 + L * q) * NZH + grp;
            v[q] = gload(srcp + ((long)(e / TB) * nx + x) * TB + (e % TB));
        }
        C *col = s + grp;
        line_dft<C, N, true>(v, tw, n1, [&](int i) { return col + i * NZH; });
        wave_lds_sync();
#pragma unroll
        for (int o = 0; o < R; ++o) col[line_out_index<N>(n1, o) * NZH] = v[o];
    }
    // E of the two rows this lane updates (pairs of points, consecutive
    // lanes on consecutive pairs), loaded once the spectrum is in: issued at
    // the start, these loads made the y pass wait for them as well (a wait
    // on a load also waits on every load issued before it)
    // (at N = 128 the 16-point z DFT needs those registers: loaded after it)
    double2 ev[2][R / 2];
    auto load_e = [&]() {
        if (mode == 0 && !creep_set && grp < N / 2) {
            const double2 *E = (const double2 *)(r.E6 + off);
#pragma unroll
            for (int q = 0; q < R / 2; ++q) {
                ev[0][q] = gload(E + n1 + L * q);
                ev[1][q] = gload(E + N / 2 + n1 + L * q);
            }
        }
    };
    if (N <= 64) load_e();
    __syncthreads();
    // z: rows 2p (a) and 2p+1 (b) as one complex row Z = A + iB, A and B
    // Hermitian-extended from their half spectra, inverse DFT along z
    if (grp < N / 2) {
        C *line = s + grp * RS;
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int k = n1 + L * q;
            C a, b;
            if (k <= N / 2) {
                a = line[k];
                b = line[NZH + k];
                if (k == 0 || 2 * k == N) { a.y = (T)0; b.y = (T)0; }
            } else {
                a = line[N - k];
                b = line[NZH + N - k];
                a.y = -a.y;
                b.y = -b.y;
            }
            v[q] = C{a.x - b.y, a.y + b.x};
        }
        line_dft<C, N, true>(v, tw, n1, [&](int i) { return line + i; });
        if (N > 64) load_e();
        // back through the line's slots to pairs of consecutive points
        wave_lds_sync();
#pragma unroll
        for (int o = 0; o < R; ++o) line[line_out_index<N>(n1, o)] = v[o];
        wave_lds_sync();
        double2 va[R / 2], vb[R / 2];
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
            const int z = 2 * (n1 + L * q);
            const C u0 = line[z], u1 = line[z + 1];
            va[q] = double2{(double)u0.x, (double)u1.x};
            vb[q] = double2{(double)u0.y, (double)u1.y};
        }
        if (mode == 0) {
            double2 *E = (double2 *)(r.E6 + off);
            const double sh = creep_set ? 0.0 : ms->shift6[c];
            const double cr = ms->creep_sym6[c];
#pragma unroll
            for (int q = 0; q < R / 2; ++q) {
                const int i = n1 + L * q;
                if (creep_set) {
                    gstore(E + i, double2{cr + (sh + va[q].x), cr + (sh + va[q].y)});
                    gstore(E + N / 2 + i, double2{cr + (sh + vb[q].x), cr + (sh + vb[q].y)});
                } else {
                    gstore(E + i, double2{ev[0][q].x + (sh + va[q].x), ev[0][q].y + (sh + va[q].y)});
                    gstore(E + N / 2 + i, double2{ev[1][q].x + (sh + vb[q].x), ev[1][q].y + (sh + vb[q].y)});
                }
            }
            // diagnostics: the running sum of the Green updates (the
            // reference's fluctuation, evolution.c:392) for the components
            // new_position reads
            const int fs = (c == 1) ? 0 : (c == 2) ? 1 : (c == 3) ? 2 : (c == 5) ? 3 : -1;
            if (r.FLU && fs >= 0) {
                double2 *F = (double2 *)(r.FLU + (long)fs * r.V + ((long)x * N + 2 * grp) * N);
#pragma unroll
                for (int q = 0; q < R / 2; ++q) {
                    const int i = n1 + L * q;
                    const double2 fa = gload(F + i), fb = gload(F + N / 2 + i);
                    gstore(F + i, double2{fa.x + va[q].x, fa.y + va[q].y});
                    gstore(F + N / 2 + i, double2{fb.x + vb[q].x, fb.y + vb[q].y});
                }
            }
        } else {
            double2 *W = (double2 *)(r.WF + off);
            double2 *F = (r.FLU && c < 2) ? (double2 *)(r.FLU + (long)(4 + c) * r.V + ((long)x * N + 2 * grp) * N)
                                          : nullptr;
#pragma unroll
            for (int q = 0; q < R / 2; ++q) {
                const int i = n1 + L * q;
                gstore(W + i, va[q]);
                gstore(W + N / 2 + i, vb[q]);
                if (F) {
                    const double2 fa = gload(F + i), fb = gload(F + N / 2 + i);
                    gstore(F + i, double2{fa.x + va[q].x, fa.y + va[q].y});
                    gstore(F + N / 2 + i, double2{fb.x + vb[q].x, fb.y + vb[q].y});
                }
            }
        }
    }
}

// ------------------------------------------------------------------
// Line-parallel split plane passes (plane variant 6) for Ny = Nz = 256:
// planes too large for one block's LDS, transformed as in variant 5 (the z
// and y transforms of a plane are two launches, the half spectrum of z row y
// at column y nzh + kz between them), but every 256-point line belongs to a
// group of 16 lanes holding 16 points each (line_dft, one exchange through
// the line's own LDS slots) instead of a block-wide Stockham transform with
// a barrier per stage.  A block of SL_LPB groups owns SL_LPB row pairs (z
// passes: blockIdx.y = chunk of row pairs), one of SLY<T>::LPB groups as
// many kz columns (y pass).
// The z passes have no block barrier: a group loads its row pair from HBM
// and writes its half spectra (or E rows) straight back.  The y pass stages
// its 256 x SLY<T>::LPB tile through LDS, consecutive lanes on consecutive
// columns of a row (a row's run of the tile spans two 128-byte column
// blocks, since nzh = 129 is odd).
// ------------------------------------------------------------------
constexpr int SL_N = 256, SL_LPB = 8, SL_NT = 16 * SL_LPB;
// y pass: kz columns per block and threads (FP64 16: 66 KB of LDS, two
// blocks per CU, against one at 32; FP32 32)
template <typename T> struct SLY { static constexpr int LPB = sizeof(T) == 8 ? 16 : 32, NT = 16 * LPB; };
template <> struct P2Plan<256> { static constexpr int L = 16, R = 16, NT = SL_NT; };
// dynamic LDS of the z passes (SL_LPB lines of 2 (N/2 + 1) slots) and of the
// y pass (SLY<T>::LPB lines of N + 1)
template <typename T> constexpr size_t sl_lds_z() { return (size_t)SL_LPB * (SL_N + 2) * sizeof(typename cplx_of<T>::t); }
template <typename T> constexpr size_t sl_lds_y() { return (size_t)SLY<T>::LPB * (SL_N + 1) * sizeof(typename cplx_of<T>::t); }

// forward z pass: rows 2p, 2p+1 of S6 as one complex row, z DFT, split into
// half spectra, out to columns y nzh + k; the extra block lb == 6 nx
// finalises the previous iteration
template <typename T>
__global__ __launch_bounds__(SL_NT) void k_zfwd_l(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    using C = typename cplx_of<T>::t;
    constexpr int N = SL_N, L = P2Plan<N>::L, R = P2Plan<N>::R;
    constexpr int NZH = N / 2 + 1, RS = 2 * NZH, KT = N / (2 * L);
    RVE_PROLOGUE(lb0);
    const int lb = fin_first(lb0, r.nx);
    if (lb == 6 * r.nx) {
        if (blockIdx.y == 0) finalize_block<SL_NT>(r);
        return;
    }
    if (!iter_active(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx;
    const int c = lb / nx, x = lb - c * nx;
    const int grp = (int)threadIdx.x / L, n1 = (int)threadIdx.x % L;
    const int p = (int)blockIdx.y * SL_LPB + grp;   // row pair (N / 2 of them)
    C tw[R];
    {
        const C *gtw = (const C *)r.tw_z;
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) tw[k1] = gload(gtw + n1 * k1);
    }
    // the two rows in pairs of points (whole cache lines per group), into
    // the line's slots in natural order
    const double2 *src = (const double2 *)(r.S6 + (long)c * r.V + ((long)x * N + 2 * p) * N);
    C *line = s + grp * RS;
    double2 pa[R / 2], pb[R / 2];
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
        pa[q] = gload(src + n1 + L * q);
        pb[q] = gload(src + N / 2 + n1 + L * q);
    }
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
        const int z = 2 * (n1 + L * q);
        line[z] = C{(T)pa[q].x, (T)pb[q].x};
        line[z + 1] = C{(T)pa[q].y, (T)pb[q].y};
    }
    wave_lds_sync();
    C v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = line[n1 + L * q];
    line_dft<C, N, false>(v, tw, n1, [&](int i) { return line + i; });
    wave_lds_sync();
#pragma unroll
    for (int o = 0; o < R; ++o) line[line_out_index<N>(n1, o)] = v[o];
    wave_lds_sync();
    // A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i, k <= N/2
    C *dst = (C *)r.spec + (long)c * r.nspec;
#pragma unroll
    for (int t = 0; t <= KT; ++t) {
        const int k = n1 + L * t;
        if (k <= N / 2) {
            const C z0 = line[k], z1 = line[(N - k) % N];
            gstore(dst + spec_at<C>(2 * p * NZH + k, x, nx), C{(T)0.5 * (z0.x + z1.x), (T)0.5 * (z0.y - z1.y)});
            gstore(dst + spec_at<C>((2 * p + 1) * NZH + k, x, nx), C{(T)0.5 * (z0.y + z1.y), (T)0.5 * (z1.x - z0.x)});
        }
    }
}

// y pass, forward or inverse, in place on SLY<T>::LPB kz columns of one plane
// (a row's columns of a tile straddle the 128-byte column blocks, nzh = 129
// being odd: the wider the tile, the fewer blocks share a column block)
template <typename T, bool INV>
__global__ __launch_bounds__(SLY<T>::NT) void k_y_l(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap, int mode)
{
    using C = typename cplx_of<T>::t;
    constexpr int N = SL_N, L = P2Plan<N>::L, R = P2Plan<N>::R;
    constexpr int NZH = N / 2 + 1, LS = N + 1;   // LDS line stride: column tiles without bank conflicts
    RVE_PROLOGUE(lb);
    if (mode == 0 ? !iter_active(r.ms) : !rve_need_skew(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx;
    const int c = lb / nx, x = lb - c * nx;
    const int grp = (int)threadIdx.x / L, n1 = (int)threadIdx.x % L;
    const int k0 = (int)blockIdx.y * SLY<T>::LPB;
    if (k0 >= NZH) return;
    const int nk = min(SLY<T>::LPB, NZH - k0);
    C *sp = (C *)r.spec + (long)c * r.nspec;
    C tw[R];
    {
        const C *gtw = (const C *)r.tw_y;
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) tw[k1] = gload(gtw + n1 * k1);
    }
    // element e = y SLY<T>::LPB + j: consecutive lanes on consecutive columns
    constexpr int EPT = N * SLY<T>::LPB / SLY<T>::NT;
    {
        C a[EPT];
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = (int)threadIdx.x + i * SLY<T>::NT, y = e / SLY<T>::LPB, j = e % SLY<T>::LPB;
            if (j < nk) a[i] = gload(sp + spec_at<C>(y * NZH + k0 + j, x, nx));
        }
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = (int)threadIdx.x + i * SLY<T>::NT, y = e / SLY<T>::LPB, j = e % SLY<T>::LPB;
            if (j < nk) s[j * LS + y] = a[i];
        }
    }
    __syncthreads();
    if (grp < nk) {
        C *col = s + grp * LS;
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = col[n1 + L * q];
        line_dft<C, N, INV>(v, tw, n1, [&](int i) { return col + i; });
        wave_lds_sync();
#pragma unroll
        for (int o = 0; o < R; ++o) col[line_out_index<N>(n1, o)] = v[o];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int e = (int)threadIdx.x + i * SLY<T>::NT, y = e / SLY<T>::LPB, j = e % SLY<T>::LPB;
        if (j < nk) gstore(sp + spec_at<C>(y * NZH + k0 + j, x, nx), s[j * LS + y]);
    }
}

// inverse z pass: half spectra of rows 2p (A) and 2p+1 (B), Hermitian
// extended into Z = A + iB, inverse z DFT; mode 0: E6[c] += shift[c] +
// result (creep: E6[c] = creep[c] + result); mode 1: WF[c] = result; the
// fluctuation sums as in k_plane_inv_p2
template <typename T>
__global__ __launch_bounds__(SL_NT) void k_zinv_l(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap, int mode)
{
    using C = typename cplx_of<T>::t;
    constexpr int N = SL_N, L = P2Plan<N>::L, R = P2Plan<N>::R;
    constexpr int NZH = N / 2 + 1;
    static_assert(L == R, "lane n1 holds z = n1 + L o after line_dft");
    RVE_PROLOGUE(lb);
    if (mode == 0 ? !iter_active(r.ms) : !rve_need_skew(r.ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int nx = r.nx;
    const int c = lb / nx, x = lb - c * nx;
    const int grp = (int)threadIdx.x / L, n1 = (int)threadIdx.x % L;
    const int p = (int)blockIdx.y * SL_LPB + grp;
    const cptr<MacroState> ms = as_const(r.ms);
    const bool creep_set = mode == 0 && (ms->creep != 0) && (ms->iter == 0);
    C tw[R];
    {
        const C *gtw = (const C *)r.tw_z;
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) tw[k1] = gload(gtw + n1 * k1);
    }
    const C *sp = (const C *)r.spec + (long)c * r.nspec;
    C v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int k = n1 + L * q, kk = (k <= N / 2) ? k : N - k;
        C a = gload(sp + spec_at<C>(2 * p * NZH + kk, x, nx));
        C b = gload(sp + spec_at<C>((2 * p + 1) * NZH + kk, x, nx));
        if (k == 0 || 2 * k == N) { a.y = (T)0; b.y = (T)0; }
        if (k > N / 2) { a.y = -a.y; b.y = -b.y; }
        v[q] = C{a.x - b.y, a.y + b.x};
    }
    C *line = s + grp * N;
    line_dft<C, N, true>(v, tw, n1, [&](int i) { return line + i; });
    // lane n1 holds z = n1 + L o in v[o]: real part row 2p, imaginary part row 2p+1
    const long off = (long)c * r.V + ((long)x * N + 2 * p) * N;
    if (mode == 0) {
        const gmut<double> E = as_global_mut(r.E6 + off);
        const double sh = creep_set ? 0.0 : ms->shift6[c];
        const double cr = ms->creep_sym6[c];
        double ea[R], eb[R];
        // E is loaded after the DFT, whose 16-point registers the loads
        // would otherwise hold
        __asm__ volatile("" ::: "memory");
        if (!creep_set) {
#pragma unroll
            for (int o = 0; o < R; ++o) {
                ea[o] = E[n1 + L * o];
                eb[o] = E[N + n1 + L * o];
            }
        }
#pragma unroll
        for (int o = 0; o < R; ++o) {
            const int z = n1 + L * o;
            const double va = (double)v[o].x, vb = (double)v[o].y;
            E[z] = creep_set ? cr + (sh + va) : ea[o] + (sh + va);
            E[N + z] = creep_set ? cr + (sh + vb) : eb[o] + (sh + vb);
        }
        const int fs = (c == 1) ? 0 : (c == 2) ? 1 : (c == 3) ? 2 : (c == 5) ? 3 : -1;
        if (r.FLU && fs >= 0) {
            const gmut<double> F = as_global_mut(r.FLU + (long)fs * r.V + ((long)x * N + 2 * p) * N);
#pragma unroll
            for (int o = 0; o < R; ++o) {
                const int z = n1 + L * o;
                F[z] += (double)v[o].x;
                F[N + z] += (double)v[o].y;
            }
        }
    } else {
        const gmut<double> W = as_global_mut(r.WF + off);
#pragma unroll
        for (int o = 0; o < R; ++o) {
            const int z = n1 + L * o;
            W[z] = (double)v[o].x;
            W[N + z] = (double)v[o].y;
        }
        if (r.FLU && c < 2) {
            const gmut<double> F = as_global_mut(r.FLU + (long)(4 + c) * r.V + ((long)x * N + 2 * p) * N);
#pragma unroll
            for (int o = 0; o < R; ++o) {
                const int z = n1 + L * o;
                F[z] += (double)v[o].x;
                F[N + z] += (double)v[o].y;
            }
        }
    }
}

// unit frequency vector, init.c:660-677
template <typename T>
__device__ __forceinline__ void kvec(int kx, int ky, int kz, int nx, int ny, int nz, T n[3])
{
    double t0 = (double)kx * (double)(1.0 / nx);
    double t1 = (double)ky * (double)(1.0 / ny);
    double t2 = (double)kz * (double)(1.0 / nz);
    t0 = (t0 - round(t0)) * TWO_PI_REF;
    t1 = (t1 - round(t1)) * TWO_PI_REF;
    t2 = (t2 - round(t2)) * TWO_PI_REF;
    const double g = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    const double ig = 1. / g;
    n[0] = (T)(t0 * ig);
    n[1] = (T)(t1 * ig);
    n[2] = (T)(t2 * ig);
}

// Green operator applied to one k point (init.c:595-610, evolution.c:363-377):
// u = -omega(n) (sigma_hat n) / V with omega = (C0_ijkl n_j n_l)^-1 by
// cofactors; d = B-basis components of sym(u (x) n); w = skew(u (x) n)
// components (2,1), (0,2), (1,0).
// Coefficients of K(n) = C0_ijkl n_j n_l as a quadratic form: K_q =
// sum_p G[6q + p] m_p for the 6 independent entries q = (00, 11, 22, 12, 02,
// 01) and m = (n0^2, n1^2, n2^2, n1 n2, n0 n2, n0 n1) (host-computed, in the
// FFT type of the kernel).
template <typename T> __device__ __forceinline__ cptr<T> green_coef(const MacroState *ms);
template <> __device__ __forceinline__ cptr<double> green_coef<double>(const MacroState *ms) { return as_const(ms->G36); }
template <> __device__ __forceinline__ cptr<float> green_coef<float>(const MacroState *ms) { return as_const(ms->G36f); }

template <typename T, typename C>
__device__ __forceinline__ void green_apply(const cptr<T> G, const T n[3], const C a[6], T invV, C d[6], C w[3])
{
    const T m[6] = {n[0] * n[0], n[1] * n[1], n[2] * n[2], n[1] * n[2], n[0] * n[2], n[0] * n[1]};
    T Kq[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        T acc = 0;
#pragma unroll
        for (int p = 0; p < 6; ++p) acc += G[6 * q + p] * m[p];
        Kq[q] = acc;
    }
    const T K[3][3] = {{Kq[0], Kq[5], Kq[4]}, {Kq[5], Kq[1], Kq[3]}, {Kq[4], Kq[3], Kq[2]}};
    const T det = K[0][0] * (K[1][1] * K[2][2] - K[1][2] * K[2][1]) - K[0][1] * (K[1][0] * K[2][2] - K[1][2] * K[2][0]) +
                  K[0][2] * (K[1][0] * K[2][1] - K[2][0] * K[1][1]);
    const T id = (T)1 / det;
    T om[3][3];
    om[0][0] = (K[1][1] * K[2][2] - K[1][2] * K[2][1]) * id;
    om[0][1] = (K[0][2] * K[2][1] - K[0][1] * K[2][2]) * id;
    om[0][2] = (K[0][1] * K[1][2] - K[0][2] * K[1][1]) * id;
    om[1][0] = (K[1][2] * K[2][0] - K[1][0] * K[2][2]) * id;
    om[1][1] = (K[0][0] * K[2][2] - K[0][2] * K[2][0]) * id;
    om[1][2] = (K[0][2] * K[1][0] - K[0][0] * K[1][2]) * id;
    om[2][0] = (K[1][0] * K[2][1] - K[1][1] * K[2][0]) * id;
    om[2][1] = (K[0][1] * K[2][0] - K[0][0] * K[2][1]) * id;
    om[2][2] = (K[0][0] * K[1][1] - K[0][1] * K[1][0]) * id;
    const T q2 = (T)RSQ2, q3 = (T)RSQ3, q6 = (T)RSQ6;
    C sg[3][3];
    sg[0][0] = C{-q2 * a[0].x - q6 * a[1].x + q3 * a[5].x, -q2 * a[0].y - q6 * a[1].y + q3 * a[5].y};
    sg[1][1] = C{q2 * a[0].x - q6 * a[1].x + q3 * a[5].x, q2 * a[0].y - q6 * a[1].y + q3 * a[5].y};
    sg[2][2] = C{(T)2 * q6 * a[1].x + q3 * a[5].x, (T)2 * q6 * a[1].y + q3 * a[5].y};
    sg[1][2] = sg[2][1] = C{q2 * a[2].x, q2 * a[2].y};
    sg[0][2] = sg[2][0] = C{q2 * a[3].x, q2 * a[3].y};
    sg[0][1] = sg[1][0] = C{q2 * a[4].x, q2 * a[4].y};
    C tv[3], u[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        tv[i] = C{(sg[i][0].x * n[0] + sg[i][1].x * n[1] + sg[i][2].x * n[2]) * invV,
                  (sg[i][0].y * n[0] + sg[i][1].y * n[1] + sg[i][2].y * n[2]) * invV};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        u[i] = C{-(om[i][0] * tv[0].x + om[i][1] * tv[1].x + om[i][2] * tv[2].x),
                 -(om[i][0] * tv[0].y + om[i][1] * tv[1].y + om[i][2] * tv[2].y)};
#define UN(i, j) C{u[i].x * n[j], u[i].y * n[j]}
    const C u00 = UN(0, 0), u11 = UN(1, 1), u22 = UN(2, 2);
    const C u12 = UN(1, 2), u21 = UN(2, 1), u02 = UN(0, 2), u20 = UN(2, 0), u01 = UN(0, 1), u10 = UN(1, 0);
#undef UN
    d[0] = C{q2 * (u11.x - u00.x), q2 * (u11.y - u00.y)};
    d[1] = C{q6 * ((T)2 * u22.x - u00.x - u11.x), q6 * ((T)2 * u22.y - u00.y - u11.y)};
    d[2] = C{q2 * (u12.x + u21.x), q2 * (u12.y + u21.y)};
    d[3] = C{q2 * (u02.x + u20.x), q2 * (u02.y + u20.y)};
    d[4] = C{q2 * (u01.x + u10.x), q2 * (u01.y + u10.y)};
    d[5] = C{q3 * (u00.x + u11.x + u22.x), q3 * (u00.y + u11.y + u22.y)};
    w[0] = C{(T)0.5 * (u21.x - u12.x), (T)0.5 * (u21.y - u12.y)};
    w[1] = C{(T)0.5 * (u02.x - u20.x), (T)0.5 * (u02.y - u20.y)};
    w[2] = C{(T)0.5 * (u10.x - u01.x), (T)0.5 * (u10.y - u01.y)};
}

// x-pass tile staging.  Element e = (x, t) of a tile of tk (a power of two,
// 2^lgt) columns starting at col0, ncol of them valid; component cc of the
// column-blocked spectrum at cc * cstr + ((col / TB) * nx + x) * TB + col % TB,
// in LDS at line (cc * tk + t), padded to nx + 1.  All NC components of an
// element are loaded before any is stored.
template <int NC, int NT, typename C>
__device__ __forceinline__ void xtile_in(const C *__restrict__ g, C *__restrict__ l, long cstr, int nx, int tk,
                                         int lgt, int ncol, int col0)
{
    constexpr int TB = 128 / (int)sizeof(C);
    const int per = nx << lgt, NXP = nx + 1;
    for (int e = (int)threadIdx.x; e < per; e += NT) {
        const int x = e >> lgt, t = e & (tk - 1);
        if (t < ncol) {
            const int col = col0 + t;
            const long o = ((long)(col / TB) * nx + x) * TB + (col % TB);
            C v[NC];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) v[cc] = g[o + cc * cstr];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) l[(cc * tk + t) * NXP + x] = v[cc];
        }
    }
}

template <int NC, int NT, typename C>
__device__ __forceinline__ void xtile_out(const C *__restrict__ l, C *__restrict__ g, long cstr, int nx, int tk,
                                          int lgt, int ncol, int col0)
{
    constexpr int TB = 128 / (int)sizeof(C);
    const int per = nx << lgt, NXP = nx + 1;
    for (int e = (int)threadIdx.x; e < per; e += NT) {
        const int x = e >> lgt, t = e & (tk - 1);
        if (t < ncol) {
            const int col = col0 + t;
            const long o = ((long)(col / TB) * nx + x) * TB + (col % TB);
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) g[o + cc * cstr] = l[(cc * tk + t) * NXP + x];
        }
    }
}

// ------------------------------------------------------------------
// x-direction FFT + Green operator.  grid ceil(ny*nzh / tk)
// mode 0 (iteration): SPEC (6 comps) -> fwd x FFT -> Delta = B(sym(u (x) n))
//   with u = -omega(n) sigma_hat n / V -> inverse x FFT -> SPEC;
//   USUM += skew(u (x) n).
// mode 1 (skew): USUM (3 comps, already in k space) -> inverse x FFT -> SPEC[0..2]
// ------------------------------------------------------------------
template <typename T, int NT, int MAXEL>
__global__ __launch_bounds__(NT) void k_xgreen(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap, int mode)
{
    using C = typename cplx_of<T>::t;
    RVE_PROLOGUE(lb);
    const int tk = r.tk;
    const MacroState *ms = r.ms;
    if (mode == 0 ? !iter_active(ms) : !rve_need_skew(ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    // Spectra and the skew sums are stored column-blocked: component cc,
    // plane x, column col = ky * nzh + kz at ((cc * ncb + col / TB) * nx + x)
    // * TB + col % TB, TB columns = 128 bytes, so a block's tile is read and
    // written in whole cache lines.  In LDS every line (one component, one
    // column, all x) is padded to nx + 1 elements against bank conflicts.
    constexpr int TB = 128 / (int)sizeof(C);
    const int nx = r.nx, ny = r.ny, nz = r.nz, nzh = r.nzh;
    const int ncolt = ny * nzh, ncb = r.ncb, NXP = nx + 1;
    const int ncomp = (mode == 0) ? 6 : 3;
    C *s = (C *)smem;
    C *u = s + 6 * tk * NXP;                        // skew sums (mode 0)
    C *twx = u + (mode == 0 ? 3 * tk * NXP : 0);
    const int tid = threadIdx.x;
    const C *gtx = (const C *)r.tw_x;
    for (int i = tid; i < nx; i += NT) twx[i] = gtx[i];
    // Tiles that share 128-byte lines run on one XCD, whose L2 then fetches
    // each line once: workgroups go to the 8 XCDs round robin, so local
    // blocks lb, lb + 8, ..., lb + 8 (G - 1) (one XCD) take the G = TB / tk
    // tiles of one line group.  Blocks past the last whole span of 8 G keep
    // tile = lb.
    int tile = lb;
    {
        const int G = TB / tk, span = 8 * G;
        const int nt = (ncolt + tk - 1) / tk;
        if (G > 1 && lb < (nt / span) * span) {
            const int xcd = lb & 7, j = lb >> 3;
            tile = ((j / G) * 8 + xcd) * G + (j % G);
        }
    }
    const int col0 = tile * tk;
    const int ncol = min(tk, ncolt - col0);
    const bool first = (as_const(ms)->iter == 0);
    const C *src = (const C *)((mode == 0) ? r.spec : r.usum);
    C *usum = (C *)r.usum;
    const int lgt = 31 - __builtin_clz(tk);   // tk is a power of two
    const long cstr = (long)ncb * nx * TB;     // one component (= nspec)
    if (mode == 0) {
        xtile_in<6, NT>(src, s, cstr, nx, tk, lgt, ncol, col0);
        if (!first) xtile_in<3, NT>(usum, u, cstr, nx, tk, lgt, ncol, col0);
    } else {
        xtile_in<3, NT>(src, s, cstr, nx, tk, lgt, ncol, col0);
    }
    __syncthreads();
    if (mode == 0) fft_lines<C, NT, MAXEL, false, true>(s, 6 * tk, nx, NXP, 1, twx, r.x_r8);
    const T invV = (T)(1.0 / (double)r.V);
    if (mode == 0) {
        const FastDiv bynx(nx), bynzh(nzh);
        for (int e = tid; e < nx * ncol; e += NT) {
            int x, kz;
            const int t = bynx.div(e, x);
            const int col = col0 + t;
            const int ky = bynzh.div(col, kz);
            C a[6];
#pragma unroll
            for (int cc = 0; cc < 6; ++cc) a[cc] = s[(cc * tk + t) * NXP + x];
            C d[6], w[3];
            if (x == 0 && ky == 0 && kz == 0) {
                // Gamma(k = 0) = 0 (init.c:586-591)
#pragma unroll
                for (int cc = 0; cc < 6; ++cc) d[cc] = C{0, 0};
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) w[cc] = C{0, 0};
            } else {
                T n[3];
                kvec<T>(x, ky, kz, nx, ny, nz, n);
                green_apply<T, C>(green_coef<T>(ms), n, a, invV, d, w);
                // The reference keeps the full spectrum and takes Re() of the
                // inverse (evolution.c:383-391), i.e. it applies the Hermitian
                // average of Gamma at k and at the index of -k.  With round()
                // mapping index N/2 to -pi (init.c:661-667) the two differ on
                // the x and y Nyquist planes; the z half-spectrum stores only
                // one of k, -k for 0 < kz < Nz/2, so average there explicitly
                // (kz = 0 and Nz/2 are averaged by the real-part extraction).
                const bool nyq = (2 * x == nx) || (2 * ky == ny);
                if (nyq && kz > 0 && 2 * kz < nz) {
                    T m[3];
                    kvec<T>((nx - x) % nx, (ny - ky) % ny, nz - kz, nx, ny, nz, m);
                    C d2[6], w2[3];
                    green_apply<T, C>(green_coef<T>(ms), m, a, invV, d2, w2);
#pragma unroll
                    for (int cc = 0; cc < 6; ++cc) d[cc] = C{(T)0.5 * (d[cc].x + d2[cc].x), (T)0.5 * (d[cc].y + d2[cc].y)};
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) w[cc] = C{(T)0.5 * (w[cc].x + w2[cc].x), (T)0.5 * (w[cc].y + w2[cc].y)};
                }
            }
#pragma unroll
            for (int cc = 0; cc < 6; ++cc) s[(cc * tk + t) * NXP + x] = d[cc];
            // sum of the skew parts over the increment (lattice spin)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                C *up = u + (i * tk + t) * NXP + x;
                *up = first ? w[i] : cadd(*up, w[i]);
            }
        }
    }
    __syncthreads();
    fft_lines<C, NT, MAXEL, true, true>(s, ncomp * tk, nx, NXP, 1, twx, r.x_r8);
    C *dst = (C *)r.spec;
    if (mode == 0) {
        xtile_out<6, NT>(s, dst, cstr, nx, tk, lgt, ncol, col0);
        xtile_out<3, NT>(u, usum, cstr, nx, tk, lgt, ncol, col0);
    } else {
        xtile_out<3, NT>(s, dst, cstr, nx, tk, lgt, ncol, col0);
    }
}

// ------------------------------------------------------------------
// Line-parallel x pass for Nx = 16, 32, 64 (x-pass variant 3): one wave
// owns all six component lines of CPW columns (CPW = 1 for Nx = 64, 2 below):
// groups of L lanes run the forward x DFTs (line_dft), then every lane
// applies the Green operator to one k point of the wave's columns (the six
// components meet in the wave's LDS), then the groups run the inverse x DFTs.
// No block barrier: the waves of a block (WPB of them, CPW * WPB columns,
// the tile tk) are independent; the skew sums of the k points are loaded
// at the start.  Mode 1 runs the inverse x DFT of the three skew sums.
// ------------------------------------------------------------------
template <int NX> struct XP2 {
    static constexpr int L = P2Plan<NX>::L, CPW = (NX == 64) ? 1 : 2;
};

template <typename T, int NX, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_xgreen_p2(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap,
                                                        int mode)
{
    using C = typename cplx_of<T>::t;
    constexpr int L = XP2<NX>::L, R = P2Plan<NX>::R, CPW = XP2<NX>::CPW, CPB = CPW * WPB;
    constexpr int TB = 128 / (int)sizeof(C);
    RVE_PROLOGUE(lb);
    const MacroState *ms = r.ms;
    if (mode == 0 ? !iter_active(ms) : !rve_need_skew(ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int ny = r.ny, nz = r.nz, nzh = r.nzh;
    const int ncolt = ny * nzh;
    // tiles sharing 128-byte lines on one XCD (as in k_xgreen)
    int tile = lb;
    {
        const int G = TB / CPB, span = 8 * G;
        const int nt = (ncolt + CPB - 1) / CPB;
        if (G > 1 && lb < (nt / span) * span) {
            const int xcd = lb & 7, j = lb >> 3;
            tile = ((j / G) * 8 + xcd) * G + (j % G);
        }
    }
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int wc0 = tile * CPB + w * CPW;   // the wave's first column
    C *ws = (C *)smem + w * (6 * CPW * NX);  // line (j, cc) of the wave at ws + (j * 6 + cc) * NX
    const long cstr = (long)r.ncb * NX * TB;
    auto gidx = [&](int col, int x) { return ((long)(col / TB) * NX + x) * TB + (col % TB); };
    // DFT role: group g of L lanes holds component cc of column j
    const int g = lane / L, n1 = lane % L;
    const int cc = g % 6, j = g / 6, col = wc0 + j;
    const bool dft = (g < 6 * CPW) && (col < ncolt) && (mode == 0 || cc < 3);
    // Green role: lane p takes k point (column j2, kx)
    const int j2 = lane / NX, kx = lane % NX, col2 = wc0 + j2;
    const bool pt = (lane < CPW * NX) && (col2 < ncolt) && mode == 0;
    const bool first = (as_const(ms)->iter == 0);
    C *usum = (C *)r.usum;
    C u[3];
    if (pt && !first) {
#pragma unroll
        for (int i = 0; i < 3; ++i) u[i] = gload(usum + i * cstr + gidx(col2, kx));
    }
    C tw[R];
    {
        const C *gtw = (const C *)r.tw_x;
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) tw[k1] = gload(gtw + n1 * k1);
    }
    C *line = ws + (j * 6 + cc) * NX;
    C *dst = (C *)r.spec;
    if (mode == 1) {
        // the skew sums back to x: inverse x DFT of USUM into SPEC[0..2]
        if (dft) {
            C v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) v[q] = gload(usum + cc * cstr + gidx(col, n1 + L * q));
            line_dft<C, NX, true>(v, tw, n1, [&](int i) { return line + i; });
#pragma unroll
            for (int o = 0; o < R; ++o) gstore(dst + cc * cstr + gidx(col, line_out_index<NX>(n1, o)), v[o]);
        }
        return;
    }
    if (dft) {
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = gload(dst + cc * cstr + gidx(col, n1 + L * q));
        line_dft<C, NX, false>(v, tw, n1, [&](int i) { return line + i; });
        wave_lds_sync();
#pragma unroll
        for (int o = 0; o < R; ++o) line[line_out_index<NX>(n1, o)] = v[o];
    }
    wave_lds_sync();
    if (pt) {
        const int ky = col2 / nzh, kz = col2 - ky * nzh;
        C *pl = ws + j2 * 6 * NX + kx;   // component c at pl[c * NX]
        C a[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) a[c] = pl[c * NX];
        C d[6], wk[3];
        if (kx == 0 && ky == 0 && kz == 0) {
            // Gamma(k = 0) = 0 (init.c:586-591)
#pragma unroll
            for (int c = 0; c < 6; ++c) d[c] = C{0, 0};
#pragma unroll
            for (int c = 0; c < 3; ++c) wk[c] = C{0, 0};
        } else {
            const T invV = (T)(1.0 / (double)r.V);
            T n[3];
            kvec<T>(kx, ky, kz, NX, ny, nz, n);
            green_apply<T, C>(green_coef<T>(ms), n, a, invV, d, wk);
            // Hermitian average on the x and y Nyquist planes, as in k_xgreen
            const bool nyq = (2 * kx == NX) || (2 * ky == ny);
            if (nyq && kz > 0 && 2 * kz < nz) {
                T m[3];
                kvec<T>((NX - kx) % NX, (ny - ky) % ny, nz - kz, NX, ny, nz, m);
                C d2[6], w2[3];
                green_apply<T, C>(green_coef<T>(ms), m, a, invV, d2, w2);
#pragma unroll
                for (int c = 0; c < 6; ++c) d[c] = C{(T)0.5 * (d[c].x + d2[c].x), (T)0.5 * (d[c].y + d2[c].y)};
#pragma unroll
                for (int c = 0; c < 3; ++c) wk[c] = C{(T)0.5 * (wk[c].x + w2[c].x), (T)0.5 * (wk[c].y + w2[c].y)};
            }
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) pl[c * NX] = d[c];
        // sum of the skew parts over the increment (lattice spin)
#pragma unroll
        for (int i = 0; i < 3; ++i) gstore(usum + i * cstr + gidx(col2, kx), first ? wk[i] : cadd(u[i], wk[i]));
    }
    wave_lds_sync();
    if (dft) {
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = line[n1 + L * q];
        line_dft<C, NX, true>(v, tw, n1, [&](int i) { return line + i; });
#pragma unroll
        for (int o = 0; o < R; ++o) gstore(dst + cc * cstr + gidx(col, line_out_index<NX>(n1, o)), v[o]);
    }
}

// ------------------------------------------------------------------
// x pass staged through LDS (x-pass variant 4, Nx = 16, 32, 48, 64, 128 or
// 256).  A block
// owns TK columns of one 128-byte column block (all of its TB columns, a
// half or a quarter).  For each component the block's elements (kx, column)
// form one contiguous run of Nx * TK complex values, so the spectrum and
// skew-sum loads and stores of a lane are those of an element-wise kernel
// (a wave touches whole lines), where variant 3 reads each line once per
// column.  Phases, separated by block barriers: load -> forward line DFTs
// (waves as in variant 3: groups of L lanes per line) -> Green operator per
// k point -> inverse line DFTs -> store.  A wave holds one column's six
// lines from Nx = 64 on (two columns' below, four lines of the block's 6 TK
// at 256, whose 256-point lines take 16 lanes), so at Nx = 128 a thread takes
// two k points, (kx, j) and (kx + 64, j), at Nx = 48 one or two, (kx, j)
// and, for kx < 16, (kx + 32, j), and at 256 up to three, 96 apart.
//
// In LDS, element (c, i, j) (component c, line index i, column j of the
// block) has the row-major index a = (c Nx + i) TK + j, moved within its
// 128-byte chunk of TB elements to a ^ f(a / TB), with
// f(m) = (m ^ m >> s1 ^ m >> s2 ^ m >> s3) mod TB (a zero shift is no term;
// SW packs s1, s2, s3 as hex digits).  The shifts, per element size, Nx and
// TK, are ones under which the lanes that one LDS cycle serves (the DFT
// groups' strided line accesses and the point-wise accesses) fall in
// distinct banks, 2 ways at most for FP32 at Nx = 16, 32; they were found by
// search over all the kernel's access patterns with 32 banks of 4 bytes.
// ------------------------------------------------------------------
template <typename T, int NX, int TK> struct XC {
    using C = typename cplx_of<T>::t;
    static constexpr int TB = 128 / (int)sizeof(C);                 // columns per 128-byte block
    static constexpr int L = P2Plan<NX>::L, R = P2Plan<NX>::R;
    // DFT role: a wave holds LPW of the block's 6 TK lines (whole columns'
    // six lines when its 64 / L groups allow, else 64 / L lines)
    static constexpr int GPW = 64 / L, LPW = GPW >= 6 ? (GPW / 6) * 6 : GPW;
    static constexpr int WPB = (6 * TK) / LPW, NT = 64 * WPB;
    static constexpr int KST = NT / TK;                              // kx step between a thread's k points
    static constexpr int PPT = (NX + KST - 1) / KST;                 // k points per thread (at most)
    static_assert(TK == TB || 2 * TK == TB || 4 * TK == TB, "a block holds a column block, a half or a quarter");
    static_assert((6 * TK) % LPW == 0, "whole waves of lines");
    static_assert(NT <= 1024, "threads per block");
    // the swizzle shifts s1 s2 s3 as hex digits (0: no term)
    static constexpr int SW = sizeof(C) == 16
        ? (NX == 16 ? (TK == 4 ? 0x012 : 0x002) : NX == 32 ? 0x013 : (TK == 8 ? 0x003 : TK == 4 ? 0x002 : 0x001))
        : (NX == 16 ? (TK == 8 ? 0x001 : 0x012) : NX == 32 ? (TK == 16 ? 0x013 : TK == 8 ? 0x136 : 0x023) : (TK == 16 ? 0x003 : 0x013));
    __device__ __forceinline__ static int shr(int m, int k) { return k ? (m >> k) : 0; }
    __device__ __forceinline__ static int slot(int c, int i, int j)
    {
        const int a = (c * NX + i) * TK + j, m = a / TB;
        return a ^ ((m ^ shr(m, SW >> 8) ^ shr(m, (SW >> 4) & 15) ^ shr(m, SW & 15)) & (TB - 1));
    }
};

template <typename T, int NX, int TK>
__global__ __launch_bounds__((XC<T, NX, TK>::NT)) void k_xgreen_c(const RveDev *__restrict__ rtab,
                                                               const int2 *__restrict__ bmap, int mode)
{
    using X = XC<T, NX, TK>;
    using C = typename X::C;
    constexpr int TB = X::TB, L = X::L, R = X::R;
    RVE_PROLOGUE(lb);
    const MacroState *ms = r.ms;
    if (mode == 0 ? !iter_active(ms) : !rve_need_skew(ms)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    C *s = (C *)smem;
    const int ny = r.ny, nz = r.nz, nzh = r.nzh;
    const int ncolt = ny * nzh;
    // the two halves of a column block run on one XCD (one L2 fetches its lines)
    int tile = lb;
    {
        constexpr int G = TB / TK, span = 8 * G;
        const int nt = (ncolt + TK - 1) / TK;
        if (G > 1 && lb < (nt / span) * span) {
            const int xcd = lb & 7, j = lb >> 3;
            tile = ((j / G) * 8 + xcd) * G + (j % G);
        }
    }
    const int col0 = tile * TK;                              // the block's first column
    const long cstr = (long)r.ncb * NX * TB;
    // element (kx, col0 + j) of a component at base + kx * TB + j
    const long base = (long)(col0 / TB) * NX * TB + (col0 % TB);
    // point role: thread p takes the k points (pkx + h KST, column pj),
    // h < PPT, at offsets pg + h KST TB of a component (at Nx = 48 the
    // second only for pkx < 16: hv(h))
    constexpr int PPT = X::PPT, KST = X::KST;
    const int p = (int)threadIdx.x, pkx = p / TK, pj = p % TK;
    const bool pt = (pkx < NX) && (col0 + pj < ncolt);
    const long pg = base + (long)pkx * TB + pj;
    auto hv = [&](int h) { return NX % KST == 0 || pkx + h * KST < NX; };
    // DFT role: group g of L lanes of wave w holds line li = w LPW + g,
    // component cc of column dj
    constexpr int LPW = X::LPW;
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int g = lane / L, n1 = lane % L;
    // (whole columns per wave: the same as cc = g % 6, dj = w LPW / 6 + g / 6)
    const int li = w * LPW + g;
    const int cc = LPW % 6 == 0 ? g % 6 : li % 6, dj = LPW % 6 == 0 ? w * (LPW / 6) + g / 6 : li / 6;
    const bool dft = (g < LPW) && (col0 + dj < ncolt) && (mode == 0 || cc < 3);
    C *spec = (C *)r.spec;
    C *usum = (C *)r.usum;
    const bool first = (as_const(ms)->iter == 0);
    // the twiddles first: a wait on a load also waits on every load issued
    // before it, so the DFTs then wait for the spectra only
    C tw[R];
    const C *gtw = (const C *)r.tw_x;
    auto load_tw = [&]() {
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) tw[k1] = gload(gtw + n1 * k1);
    };
    load_tw();
    // load
    C u[PPT][3];
    auto load_u = [&]() {
#pragma unroll
        for (int h = 0; h < PPT; ++h)
            if (hv(h)) {
#pragma unroll
                for (int i = 0; i < 3; ++i) u[h][i] = gload(usum + i * cstr + pg + h * KST * TB);
            }
    };
    if (pt) {
        if (mode == 0) {
            C a[PPT][6];
#pragma unroll
            for (int h = 0; h < PPT; ++h)
                if (hv(h)) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) a[h][c] = gload(spec + c * cstr + pg + h * KST * TB);
                }
            // the running skew sums after the spectra (kept in that order: a
            // wait on a load also waits on the loads issued before it), and
            // on the first iteration too (then not used), since a load under
            // a branch makes the waits after it count from the shorter path;
            // at Nx = 128 only after the forward DFT, whose 16-point lines
            // need those registers
            __asm__ volatile("" ::: "memory");
            if (PPT == 1) load_u();
#pragma unroll
            for (int h = 0; h < PPT; ++h)
                if (hv(h)) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) s[X::slot(c, pkx + h * KST, pj)] = a[h][c];
                }
        } else {
            C a[PPT][3];
#pragma unroll
            for (int h = 0; h < PPT; ++h)
                if (hv(h)) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) a[h][c] = gload(usum + c * cstr + pg + h * KST * TB);
                }
#pragma unroll
            for (int h = 0; h < PPT; ++h)
                if (hv(h)) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) s[X::slot(c, pkx + h * KST, pj)] = a[h][c];
                }
            // keeps this tail apart from the other branch's (a shared tail
            // would take this path's shorter wait counts)
            __asm__ volatile("" ::: "memory");
        }
    }
    auto line = [&](int i) { return s + X::slot(cc, i, dj); };
    __syncthreads();
    if (mode == 0) {
        if (dft) {
            C v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) v[q] = *line(n1 + L * q);
            line_dft<C, NX, false>(v, tw, n1, line);
            wave_lds_sync();
#pragma unroll
            for (int o = 0; o < R; ++o) *line(line_out_index<NX>(n1, o)) = v[o];
        }
        if (PPT > 1 && pt) load_u();
        __syncthreads();
#pragma unroll
        for (int h = 0; h < PPT; ++h) if (pt && hv(h)) {
            const int kx = pkx + h * KST, col = col0 + pj;
            const int ky = col / nzh, kz = col - ky * nzh;
            C a[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) a[c] = s[X::slot(c, kx, pj)];
            C d[6], wk[3];
            if (kx == 0 && ky == 0 && kz == 0) {
                // Gamma(k = 0) = 0 (init.c:586-591)
#pragma unroll
                for (int c = 0; c < 6; ++c) d[c] = C{0, 0};
#pragma unroll
                for (int c = 0; c < 3; ++c) wk[c] = C{0, 0};
            } else {
                const T invV = (T)(1.0 / (double)r.V);
                T n[3];
                kvec<T>(kx, ky, kz, NX, ny, nz, n);
                green_apply<T, C>(green_coef<T>(ms), n, a, invV, d, wk);
                // Hermitian average on the x and y Nyquist planes, as in k_xgreen
                const bool nyq = (2 * kx == NX) || (2 * ky == ny);
                if (nyq && kz > 0 && 2 * kz < nz) {
                    T m[3];
                    kvec<T>((NX - kx) % NX, (ny - ky) % ny, nz - kz, NX, ny, nz, m);
                    C d2[6], w2[3];
                    green_apply<T, C>(green_coef<T>(ms), m, a, invV, d2, w2);
#pragma unroll
                    for (int c = 0; c < 6; ++c) d[c] = C{(T)0.5 * (d[c].x + d2[c].x), (T)0.5 * (d[c].y + d2[c].y)};
#pragma unroll
                    for (int c = 0; c < 3; ++c) wk[c] = C{(T)0.5 * (wk[c].x + w2[c].x), (T)0.5 * (wk[c].y + w2[c].y)};
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) s[X::slot(c, kx, pj)] = d[c];
            // sum of the skew parts over the increment (lattice spin)
#pragma unroll
            for (int i = 0; i < 3; ++i) gstore(usum + i * cstr + pg + h * KST * TB, first ? wk[i] : cadd(u[h][i], wk[i]));
        }
        // at Nx = 128 the 16 twiddles are not held through the Green
        // operator: loaded again (from the cache) for the inverse DFT
        if (NX > 64) load_tw();
        __syncthreads();
    }
    // inverse x DFT (mode 1: of the skew sums, into SPEC[0..2])
    if (dft) {
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = *line(n1 + L * q);
        line_dft<C, NX, true>(v, tw, n1, line);
        wave_lds_sync();
#pragma unroll
        for (int o = 0; o < R; ++o) *line(line_out_index<NX>(n1, o)) = v[o];
    }
    __syncthreads();
    if (pt) {
        if (mode == 0) {
#pragma unroll
            for (int h = 0; h < PPT; ++h)
                if (hv(h)) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) gstore(spec + c * cstr + pg + h * KST * TB, s[X::slot(c, pkx + h * KST, pj)]);
                }
        } else {
#pragma unroll
            for (int h = 0; h < PPT; ++h)
                if (hv(h)) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) gstore(spec + c * cstr + pg + h * KST * TB, s[X::slot(c, pkx + h * KST, pj)]);
                }
        }
    }
}

// ------------------------------------------------------------------
// Constitutive pieces (double, registers)
// ------------------------------------------------------------------

// Rotation of the deviatoric B basis by R (crystal -> sample):
// Q[a][b] = B_a : (R B_b R^T), a, b < 5.  A crystal-frame Schmid vector p
// maps to the sample frame as Q p, so update_schmid (kinematics.c:426-459)
// is Q applied to Schm_xt, and the resolved shear is p . (Q^T sigma).
// Column b of the B-basis rotation Q of TranMat R: the B components of
// R B_b R^T, so that the components of R T R^T are Q t.
__device__ __forceinline__ void basis_rotation_col(const double R[9], int b, double q[5])
{
    // columns of R
    const double c0[3] = {R[0], R[3], R[6]}, c1[3] = {R[1], R[4], R[7]}, c2[3] = {R[2], R[5], R[8]};
    // M = R B_b R^T, symmetric: entries 00 11 22 12 02 01
    double m[6];
#define DY(u, v, i, j) ((u)[i] * (v)[j] + (v)[i] * (u)[j])
    if (b == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int i = k < 3 ? k : (k == 3 ? 1 : 0), j = k < 3 ? k : (k == 5 ? 1 : 2);
            m[k] = RSQ2 * (c1[i] * c1[j] - c0[i] * c0[j]);
        }
    } else if (b == 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int i = k < 3 ? k : (k == 3 ? 1 : 0), j = k < 3 ? k : (k == 5 ? 1 : 2);
            m[k] = RSQ6 * (2. * c2[i] * c2[j] - c0[i] * c0[j] - c1[i] * c1[j]);
        }
    } else {
        const double *u = (b == 2) ? c1 : c0;
        const double *w = (b == 4) ? c1 : c2;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int i = k < 3 ? k : (k == 3 ? 1 : 0), j = k < 3 ? k : (k == 5 ? 1 : 2);
            m[k] = RSQ2 * DY(u, w, i, j);
        }
    }
#undef DY
    q[0] = RSQ2 * (m[1] - m[0]);
    q[1] = RSQ6 * (2. * m[2] - m[0] - m[1]);
    q[2] = 2. * RSQ2 * m[3];
    q[3] = 2. * RSQ2 * m[4];
    q[4] = 2. * RSQ2 * m[5];
}

// Deviatoric B-basis components of R T R^T (TO_SAMPLE: crystal -> sample)
// or of R^T T R (sample -> crystal), for the traceless symmetric T whose
// components are a: the 3x3 products, cheaper than forming Q when one or two
// tensors are rotated.  The trace of the result is zero, so its (2,2) entry
// is taken as -(00 + 11).
template <bool TO_SAMPLE>
__device__ __forceinline__ void rot_dev(const double R[9], const double a[5], double out[5])
{
    const double d0 = RSQ2 * a[0], d1 = RSQ6 * a[1];
    const double e12 = RSQ2 * a[2], e02 = RSQ2 * a[3], e01 = RSQ2 * a[4];
    const double T[3][3] = {{-d0 - d1, e01, e02}, {e01, d0 - d1, e12}, {e02, e12, 2.0 * d1}};
    auto A = [&](int i, int j) { return TO_SAMPLE ? R[i * 3 + j] : R[j * 3 + i]; };
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[i][j] = A(i, 0) * T[0][j] + A(i, 1) * T[1][j] + A(i, 2) * T[2][j];
    auto Tp = [&](int i, int j) { return M[i][0] * A(j, 0) + M[i][1] * A(j, 1) + M[i][2] * A(j, 2); };
    const double t00 = Tp(0, 0), t11 = Tp(1, 1);
    out[0] = RSQ2 * (t11 - t00);
    out[1] = -3.0 * RSQ6 * (t00 + t11);
    out[2] = 2. * RSQ2 * Tp(1, 2);
    out[3] = 2. * RSQ2 * Tp(0, 2);
    out[4] = 2. * RSQ2 * Tp(0, 1);
}

__device__ __forceinline__ void basis_rotation(const double R[9], double Q[5][5])
{
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        double q[5];
        basis_rotation_col(R, b, q);
#pragma unroll
        for (int a = 0; a < 5; ++a) Q[a][b] = q[a];
    }
}

// Compiler-only barrier: forces loads of loop-invariant tables (material
// constants, per-orientation matrices) to be reissued after it instead of
// being hoisted into registers for the whole Newton loop.  Those loads hit
// the scalar cache or L1; holding them would cost occupancy.
#define RELOAD_BARRIER() __asm__ volatile("" ::: "memory")

// waves per SIMD the Newton kernel is compiled for (its LDS allows 3; with a
// separate backward CRSS it allows 2)
#ifndef VOCE_NR_WAVES
#define VOCE_NR_WAVES 3
#endif
// slip systems per trip of the flow-rule loop of the Newton kernel
#ifndef VOCE_NR_SYS_UNROLL
#define VOCE_NR_SYS_UNROLL 1
#endif
#define VOCE_PRAGMA_(x) _Pragma(#x)
#define VOCE_PRAGMA(x) VOCE_PRAGMA_(x)


#ifndef VOCE_IPOW_MODE
#define VOCE_IPOW_MODE 0
#endif
// x^n for 0 <= n < 64 by binary powering; n is wave-uniform
__device__ __forceinline__ double ipow(double x, int n)
{
    double r = 1.0;
#if VOCE_IPOW_MODE == 0
    // unrolled bit tests: scalar branches, no loop
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        if ((n >> b) == 0) break;
        if ((n >> b) & 1) r *= x;
        x *= x;
    }
#elif VOCE_IPOW_MODE == 1
    // branch-free: every bit, selects
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const double t = r * x;
        r = ((n >> b) & 1) ? t : r;
        x *= x;
    }
#else
    (void)n;
    const double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
    r = x16 * x2 * x;
#endif
    return r;
}

// StrainRate_eval (constitutive.c:4-72) in the crystal frame: tau_s = p_s . sx
// for the crystal-frame stress sx, ex = sum p_s gamdot_s (crystal frame; the
// sample-frame rate is Q ex).  gd / tx receive gamdot and tau_x when non-null
// (12 entries each).
__device__ __forceinline__ void flow_rule_x(const double sx[6], const PhaseConst *pc, int nsys, const double cf[12],
                                            const double cb[12], double ex[5], double *gd, double *tx)
{
#pragma unroll
    for (int a = 0; a < 5; ++a) ex[a] = 0.0;
    const double nsr = pc->nsr, g0 = pc->gam0;
    const int intn = pc->int_n;
    const int ni = (int)nsr;
#pragma unroll
    for (int s = 0; s < 12; ++s) {
        if (s < nsys) {
            const double *p = pc->px[s];
            const double rss0 = p[0] * sx[0] + p[1] * sx[1] + p[2] * sx[2] + p[3] * sx[3] + p[4] * sx[4];
            const double tc = (rss0 < 0) ? cb[s] : cf[s];
            const double rr = rss0 / tc;
            double pn1, pn;
            if (intn) {
                const double ar = fabs(rr);
                pn1 = ipow(ar, ni - 1);
                pn = pn1 * ar;
            } else {
                pn1 = fabs(pow(rr, nsr - 1.));
                pn = fabs(pow(rr, nsr));
            }
            const double rss1 = g0 * nsr * pn1 / tc;
            const double rss2 = g0 * pn * ((double)(2 * (rr > 0) - 1));
            if (gd) gd[s] = rss2;
            if (tx) tx[s] = rss1;
#pragma unroll
            for (int a = 0; a < 5; ++a) ex[a] += p[a] * rss2;
        }
    }
}


// x moved across lanes by a DPP control (both 32-bit halves); lanes whose
// row is masked off receive 0
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double x)
{
    const int lo = __double2loint(x), hi = __double2hiint(x);
    return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, ROWS, 0xf, false),
                            __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWS, 0xf, false));
}

// Sum of x over each 32-lane half of a wave, complete in lanes 31 and 63:
// DPP moves in a fixed order (quad swaps, half-row and row mirrors, a row
// broadcast), so the result does not depend on timing.
__device__ __forceinline__ double halfwave_sum_last(double x)
{
    x += dpp_f64<0xB1, 0xf>(x);    // quad_perm [1,0,3,2]
    x += dpp_f64<0x4E, 0xf>(x);    // quad_perm [2,3,0,1]
    x += dpp_f64<0x141, 0xf>(x);   // row_half_mirror
    x += dpp_f64<0x140, 0xf>(x);   // row_mirror: every lane holds its row's sum
    x += dpp_f64<0x142, 0xa>(x);   // row_bcast:15 into rows 1 and 3
    return x;
}

// Block sums of NV values per lane, stored by out[0 .. NV).  The values go
// through LDS in red (NV rows of NT, column = thread: a thread writes only
// its own column, so red may alias per-lane arrays the thread is done with),
// then a 32-lane group sums one value: lane j of the group adds entries
// j, j + 32, j + 64, ... in that order (consecutive lanes, consecutive
// doubles), and a DPP tree over the group adds the lanes.  The order is
// fixed.  DUP: out[NV .. NV + 5] repeat out[2 .. 7] (with one crystalline
// phase the phase-1 stress sums are the stress sums).
template <int NT, int NV, bool DUP = false>
__device__ __forceinline__ void block_reduce_store(const double (&vals)[NV], double (*red)[NT], double *out)
{
    static_assert(NT % 32 == 0, "whole 32-lane groups");
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][threadIdx.x] = vals[i];
    __syncthreads();
#pragma unroll
    for (int base = 0; base < NV * 32; base += NT) {
        const int t = base + (int)threadIdx.x, i = t >> 5, j = t & 31;
        double x = 0.0;
        if (i < NV) {
#pragma unroll
            for (int e = 0; e < NT / 32; ++e) x += red[i][j + 32 * e];
        }
        x = halfwave_sum_last(x);
        if (i < NV && j == 31) {
            out[i] = x;
            if (DUP && i >= 2) out[i + 6] = x;
        }
    }
}

// Fixed-order sum of nb partial rows (16 doubles each, the first NV used) by
// one block into tot[] (shared), in an order that does not depend on the
// block size: S row streams, stream s summing entries e < NE of rows s,
// s + S, s + 2 S, ... in row order, then the streams added in stream order.
// With dup (the Newton partials of one crystalline phase: entries 8..13
// repeat 2..7, the phase-1 stress sums being the stress sums) only entries
// 0..7 are summed, over S = 32 streams; otherwise all 16 over S = 16.  Logical lane l = 64 w +
// lane (256 of them, over four waves; fewer waves loop) takes entry l % NE
// of stream l / NE, so a load instruction reads whole 128-byte rows, and a
// lane keeps U loads in flight (one round trip per U rows of its stream).
// Ends with a barrier.
template <int NT, int NV>
__device__ __forceinline__ void sum_partials(const double *part, int nb, bool dup, double *tot)
{
    constexpr int NW = NT / 64, NWU = NW < 4 ? NW : 4, PASSES = 4 / NWU, U = 32;
    static_assert(NW == 1 || NW == 2 || NW >= 4, "waves per block");
    static_assert(NV <= 16, "at most 16 partial sums per row");
    const bool uni = dup && NV == 14;
    __shared__ double red[256];
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int lne = uni ? 3 : 4, S = uni ? 32 : 16;   // NE = 1 << lne
    const gptr<double> gp = as_global(part);
    if (w < NWU) {
#pragma unroll 1
        for (int t = 0; t < PASSES; ++t) {
            const int l = 64 * (w + NWU * t) + lane;
            const int st = l >> lne, e = l & ((1 << lne) - 1);
            double a = 0.0;
            int b = st;
            for (; b + S * (U - 1) < nb; b += S * U) {
                double q[U];
#pragma unroll
                for (int k = 0; k < U; ++k) q[k] = __builtin_nontemporal_load(gp + (long)(b + S * k) * 16 + e);
#pragma unroll
                for (int k = 0; k < U; ++k) a += q[k];
            }
            // the remaining rows of the stream (fewer than U) also go out
            // together (a loop with a run-time trip count waits for each load
            // before issuing the next), rows past the end read as the last
            // row and add 0: the sum is the same, in the same order
            if (b < nb) {
                double q[U];
#pragma unroll
                for (int k = 0; k < U; ++k) {
                    const int row = b + S * k;
                    q[k] = __builtin_nontemporal_load(gp + (long)(row < nb ? row : nb - 1) * 16 + e);
                }
#pragma unroll
                for (int k = 0; k < U; ++k)
                    if (b + S * k < nb) a += q[k];
            }
            red[l] = a;
        }
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int e = (uni && threadIdx.x >= 8) ? (int)threadIdx.x - 6 : (int)threadIdx.x;
        double acc = 0.0;
        for (int k = 0; k < S; ++k) acc += red[(k << lne) + e];
        tot[threadIdx.x] = acc;
    }
    __syncthreads();
}

// tensor of a B-basis 6-vector (chg_basis opt 1)
__device__ __forceinline__ void b6_to_t(const double a[6], double T[9])
{
    T[0] = -RSQ2 * a[0] - RSQ6 * a[1] + RSQ3 * a[5];
    T[4] = RSQ2 * a[0] - RSQ6 * a[1] + RSQ3 * a[5];
    T[8] = 2. * RSQ6 * a[1] + RSQ3 * a[5];
    T[5] = T[7] = RSQ2 * a[2];
    T[2] = T[6] = RSQ2 * a[3];
    T[1] = T[3] = RSQ2 * a[4];
}
// B-basis of the symmetric part of a 3x3 (chg_basis opt 2 of sym(T))
__device__ __forceinline__ void sym_to_b6(const double T[9], double a[6])
{
    const double s00 = T[0], s11 = T[4], s22 = T[8];
    const double s12 = (T[5] + T[7]) / 2., s02 = (T[2] + T[6]) / 2., s01 = (T[1] + T[3]) / 2.;
    a[0] = s00 * (-RSQ2) + s11 * RSQ2;
    a[1] = s00 * (-1. * RSQ6) + s11 * (-1. * RSQ6) + s22 * (2. * RSQ6);
    a[2] = s12 * RSQ2 + s12 * RSQ2;
    a[3] = s02 * RSQ2 + s02 * RSQ2;
    a[4] = s01 * RSQ2 + s01 * RSQ2;
    a[5] = s00 * RSQ3 + s11 * RSQ3 + s22 * RSQ3;
}

__device__ double von_mises_strain(const double t[9])
{
    // VonMises, evolution.c:243-262
    const double tr = t[0] + t[4] + t[8];
    double vm = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double d = (t[i * 3 + j] + t[j * 3 + i]) / 2. - (double)(i == j) * tr / 3.0;
            vm += d * d;
        }
    return sqrt(2. / 3. * vm);
}

// get_smacro's von Mises of a stress average (evolution.c:218-238)
__device__ double svm_of(const double S[9])
{
    double a[6];
    sym_to_b6(S, a);
    double dev[9];
    const double a5[6] = {a[0], a[1], a[2], a[3], a[4], 0.0};
    b6_to_t(a5, dev);
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) v += dev[i] * dev[i];
    return sqrt(3. / 2. * v);
}

// Finaliser of one iteration (evolution.c:405-428 and get_smacro,
// evolution.c:170-241), run by thread 0 on the summed partials.  `in` is
// the LDS copy of the macro state: its entries are read where they are
// used, so the finaliser holds few registers (it shares its kernel's
// register allocation with the forward plane blocks, whose occupancy it
// would otherwise set).
__device__ void finalize_iteration(MacroState *ms, const MacroState *in, const double *tot, double *lds_dD)
{
    const double wgt = in->wgt, wgt1 = in->wgt_over_ph1;
    // IJV (evolution.c:179) as nibbles: Voigt i -> (IJV0, IJV1), so a rolled
    // loop indexes no private array
    constexpr unsigned IJV0 = 0x001210u, IJV1 = 0x122210u;
    auto ij0 = [](int i) { return (int)((IJV0 >> (4 * i)) & 15u); };
    auto ij1 = [](int i) { return (int)((IJV1 >> (4 * i)) & 15u); };
    double SA[9], SA1[9];
    {
        double s6[6], s16[6];
        for (int i = 0; i < 6; ++i) { s6[i] = tot[2 + i]; s16[i] = tot[8 + i]; }
        b6_to_t(s6, SA);
        b6_to_t(s16, SA1);
    }
    for (int i = 0; i < 9; ++i) {
        SA[i] *= wgt;
        SA1[i] *= wgt1;
    }
    // mixed-BC correction, upper triangle only (evolution.c:199-214); the
    // rolled row loop collects dD in LDS
    double rhs[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int kk = ij0(k), ll = ij1(k);
        rhs[k] = in->Scauchy[kk * 3 + ll] - SA[kk * 3 + ll];
    }
    for (int i = 0; i < 9; ++i) lds_dD[i] = 0.0;
#pragma unroll 1
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
        if (in->velgrad_flag[i] == 0) {
            const int row = (ij0(i) * 3 + ij1(i)) * 9;
#pragma unroll
            for (int k = 0; k < 6; ++k)
                acc += in->S0[row + ij0(k) * 3 + ij1(k)] * in->stress_flag[k] * rhs[k];
        }
        lds_dD[ij0(i) * 3 + ij1(i)] = acc;
    }
    double dD[9];
    for (int i = 0; i < 9; ++i) dD[i] = lds_dD[i];
    const double s_vm = svm_of(SA), s_vm1 = svm_of(SA1);
    const double e_vm = in->e_vm;
    double Err_e = tot[1] * wgt, Err_s = tot[0] * wgt;
    if (fabs(e_vm) > 1E-5) Err_e /= e_vm;
    Err_s /= s_vm;
    const int it = in->iter;
    const int more = (it + 1 < in->iter_max) && (fmax(fabs(Err_s), fabs(Err_e)) > fabs(in->err_tol));
    for (int i = 0; i < 9; ++i) {
        ms->SigAvg[i] = SA[i];
        ms->SigAvg1[i] = SA1[i];
        ms->dD[i] = dD[i];
        ms->acum[i] = in->acum[i] + dD[i];
    }
    ms->s_vm = s_vm;
    ms->s_vm1 = s_vm1;
    ms->Err_e = Err_e;
    ms->Err_s = Err_s;
    if (it < in->log_cap) {
        double *lg = in->logs;
        const int cap = in->log_cap;
        lg[it] = Err_e;
        lg[cap + it] = Err_s;
        lg[2 * cap + it] = s_vm;
    }
    ms->iter = it + 1;
    ms->iter_total = in->iter_total + 1;
    if (more) {
        for (int i = 0; i < 9; ++i) ms->applied[i] = in->applied[i] + dD[i];
        double sh[6];
        sym_to_b6(dD, sh);
        for (int i = 0; i < 6; ++i) ms->shift6[i] = sh[i];
    }
    ms->active = more;   // read by the next kernels (the launch boundary orders it)
}

// ------------------------------------------------------------------
// Newton-Raphson stress update, UpdateStress (evolution.c:3-168).
//
// Per NR iteration (evolution.c:63-114) the residual and Jacobian of
// evolution.c:70-94 share the left factor C0 = C066:
//   R = C0 g,  g = K_c sigma + w + dt edot,   J = C0 K,  K = K_c + dt D
// with K_c = C0^-1 + S_local (per orientation, from the orientation table)
// and w = (eps_p - eps) - C0^-1 lambda, so the step J^-1 R = K^-1 g is one
// LDL^T solve of a symmetric positive definite 6x6 (see ldlt_solve6).
// The iterations run in the crystal frame of the voxel's TranMat R (fixed
// during an increment): with Q the B-basis rotation of R, every quantity of
// the step maps by Q6 = diag(Q, 1), the sample-frame Schmid vectors
// Schm_gr = Q Schm_xt (update_schmid, kinematics.c:426-459) become the
// constant Schm_xt, and K_c becomes the per-voxel Q6^T K_c Q6 that the
// close-out rewrites with R.  StrainRate_eval's sums (constitutive.c:49-64)
// enter one slip system at a time: edot += P_s gamdot_s and
// D += tau_x_s P_s (x) P_s.  Stress, the last flow-rule stress and edot are
// rotated back once, after the loop.
//
// Partials per block: [0] sum|sigma-lambda|, [1] sum|eps_tot-eps|,
// [2..7] sum sigma (B basis), [8..13] sum sigma over phase 1, summed in
// block order by k_fin_iter.  UNI: a single crystalline phase (uniform
// material constants).
// ------------------------------------------------------------------

// 1/d: hardware reciprocal estimate refined by two Newton steps
// r <- r + r (1 - d r), which square the relative error each time.
__device__ __forceinline__ double rcp_refined(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// Packed lower triangle of a symmetric 6x6: (i, j), j <= i, at i (i + 1) / 2 + j.
__device__ __forceinline__ constexpr int lt6(int i, int j) { return i * (i + 1) / 2 + j; }

// Solve K x = g for symmetric positive definite N x N K (packed lower
// triangle, factorised in place: L below the diagonal, the pivots on it) by
// LDL^T.  Returns 35 (the reference's singular-matrix code) if a pivot is not
// positive, which a positive definite matrix never gives.
template <int N>
__device__ __forceinline__ int ldlt_solve_packed(double K[N * (N + 1) / 2], const double g[N], double x[N])
{
    double iD[N];
    int bad = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double v[N];
#pragma unroll
        for (int k = 0; k < j; ++k) v[k] = K[lt6(j, k)] * K[lt6(k, k)];
        double d = K[lt6(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= K[lt6(j, k)] * v[k];
        bad |= !(d > 0.0);
        K[lt6(j, j)] = d;
        iD[j] = rcp_refined(d);
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            double t = K[lt6(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= K[lt6(i, k)] * v[k];
            K[lt6(i, j)] = t * iD[j];
        }
    }
    double y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double t = g[i];
#pragma unroll
        for (int k = 0; k < i; ++k) t -= K[lt6(i, k)] * y[k];
        y[i] = t;
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double t = y[i] * iD[i];
#pragma unroll
        for (int k = i + 1; k < N; ++k) t -= K[lt6(k, i)] * x[k];
        x[i] = t;
    }
    return bad ? 35 : 0;
}

// The deviatoric block of K_c (row-major 6x6, symmetric) in the crystal frame
// of the TranMat whose B-basis rotation is Q: Q^T K_c Q, written as its packed
// lower triangle, in tiles of 64 voxels (component q of voxel v at
// ((v / 64) 15 + q) 64 + v % 64).  The hydrostatic row and column couple by
// rounding only (cubic elasticity, checked on the host) and are not kept.
__device__ __forceinline__ double *kcv_voxel(const RveDev &r, long v) { return r.KCV + (v >> 6) * (VOCE_KCV_COMP * 64) + (v & 63); }

__device__ __forceinline__ void store_kc_crystal(const RveDev &r, long v, const double *Kc, const double Q[5][5])
{
    double *dst = kcv_voxel(r, v);
    double M[5][5];   // deviatoric block of K_c times Q
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            double t = 0.0;
#pragma unroll
            for (int b = 0; b < 5; ++b) t += Kc[a * 6 + b] * Q[b][j];
            M[a][j] = t;
        }
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < 5; ++a) t += Q[a][i] * M[a][j];
            dst[lt6(i, j) * 64] = t;
        }
}

// Forms of the power law x^(n-1), x^n of the flow rule: n = 20 (a fixed
// product chain), another integer n (binary powering), real n (pow).
enum { PM_N20 = 0, PM_INT = 1, PM_REAL = 2 };

// Newton iterations of UpdateStress (evolution.c:63-114) on one voxel, in the
// crystal frame: sig (crystal-frame stress, updated), ed (crystal-frame
// strain rate of the last flow-rule evaluation), lds_old (the stress it was
// evaluated at), lds_w (crystal-frame w), lds_ic (1 / CRSS).  Returns 0, or 35 (the reference's singular-matrix code) for a
// non-positive pivot.
template <unsigned V> struct UMask { static constexpr unsigned value = V; };

template <int NT, bool TWO, int PM, bool FCC>
__device__ __forceinline__ int newton_crystal(double sig[6], double ed[5], double glast[5], double tol2, double h6,
                                              cptr<PhaseConst> pc, int nsys, double dt, gptr<double> kcv,
                                              double (*lds_w)[NT], double (*lds_old)[NT], double (*lds_ic)[NT])
{
    const double nsr = pc->nsr, g0 = pc->gam0;
    const int ni = (int)nsr;
    // the loop runs while erral = |d sigma| / |lambda| > 1e-7, tested as
    // |d sigma|^2 > tol2 = 1e-14 |lambda|^2; it starts at erral = 1e-6
    // (evolution.c:58-63)
    bool more = true;
    int iterl = 0;
    while ((iterl < 100) && more) {
        ++iterl;
        RELOAD_BARRIER();
#pragma unroll
        for (int i = 0; i < 6; ++i) lds_old[i][threadIdx.x] = sig[i];
        // StrainRate_eval (constitutive.c:4-72) in the crystal frame: edot
        // and the symmetric 5x5 d(edot)/d(sigma), one slip system at a time
        double D[15];
#pragma unroll
        for (int a = 0; a < 5; ++a) ed[a] = 0.0;
#pragma unroll
        for (int q = 0; q < 15; ++q) D[q] = 0.0;
        // two slip systems per trip: their independent dependency chains
        // (resolved shear, power law) interleave
        auto rates = [&](const double *P, double icf, double icb, double &rss1, double &rss2) {
            // resolved shear: the 5-component dot of constitutive.c:26-27
            double rss0 = 0.0;
#pragma unroll
            for (int a = 0; a < 5; ++a) rss0 += P[a] * sig[a];
            const double itc = (rss0 < 0) ? icb : icf;
            const double rr = rss0 * itc;
            double pn1, pn;
            if (PM == PM_N20) {
                // x^19 with ipow's products in ipow's order
                const double ar = fabs(rr);
                const double x2 = ar * ar, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
                pn1 = (ar * x2) * x16;
                pn = pn1 * ar;
            } else if (PM == PM_INT) {
                const double ar = fabs(rr);
                pn1 = ipow(ar, ni - 1);
                pn = pn1 * ar;
            } else {
                pn1 = fabs(pow(rr, nsr - 1.));
                pn = fabs(pow(rr, nsr));
            }
            rss1 = g0 * nsr * pn1 * itc;
            rss2 = g0 * pn * ((double)(2 * (rr > 0) - 1));
        };
        auto accum = [&](const double *P, double rss1, double rss2) {
#pragma unroll
            for (int a = 0; a < 5; ++a) ed[a] += P[a] * rss2;
            int q = 0;
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                const double t = P[a] * rss1;
#pragma unroll
                for (int b2 = a; b2 < 5; ++b2) { D[q] += t * P[b2]; ++q; }
            }
        };
        int s = 0;
        if (FCC) {
            // The standard FCC systems (voce_fcc_schmid, host-checked): system
            // 3p + j is direction j of plane p, and over the four planes each
            // Schmid component is a constant times a Walsh row (A = ++++,
            // B = +-+-, C = ++--, E = +--+).  So the four resolved shears of
            // direction j are a Walsh transform of at most three scaled
            // stress components, and edot = sum_s P_s v_s and
            // D = sum_s P_s P_s^T u_s need only the Walsh transforms over the
            // planes of v = r^n sgn r and u = r^(n-1) / tau_c: a product of
            // two Walsh rows is a Walsh row.  The factors gam0 and n (and dt
            // for D) are the host's wk.
            constexpr double AL = 0.28867513459481287;   // 1 / (2 sqrt 3)
            const double a0 = AL * sig[0], b1 = 0.5 * sig[1], a2 = AL * sig[2], a3 = AL * sig[3], a4 = AL * sig[4];
            double Wu[3][4], Wv[3][4];
            auto dir = [&](int j, double r0, double r1, double r2, double r3) {
                const double rs[4] = {r0, r1, r2, r3};
                double u[4], vv[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    const int sy = 3 * p + j;
                    const double icf = lds_ic[sy][threadIdx.x];
                    const double icb = TWO ? lds_ic[TWO ? 12 + sy : 0][threadIdx.x] : icf;
                    const double itc = (rs[p] < 0) ? icb : icf;
                    const double rr = rs[p] * itc;
                    const double ar = fabs(rr);
                    const double x2 = ar * ar, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
                    const double pn1 = (ar * x2) * x16;
                    u[p] = pn1 * itc;
                    vv[p] = copysign(pn1 * ar, rr);
                }
                const double su = u[0] + u[1], du = u[0] - u[1], tu = u[2] + u[3], eu = u[2] - u[3];
                Wu[j][0] = su + tu;
                Wu[j][1] = du + eu;
                Wu[j][2] = su - tu;
                Wu[j][3] = du - eu;
                const double sv = vv[0] + vv[1], dv = vv[0] - vv[1], tv = vv[2] + vv[3], ev = vv[2] - vv[3];
                Wv[j][0] = sv + tv;
                Wv[j][1] = dv + ev;
                Wv[j][2] = sv - tv;
                Wv[j][3] = dv - ev;
            };
            {   // j = 0: Walsh coefficients A = a4, B = a3, E = a0 - b1
                const double E = a0 - b1, sAB = a4 + a3, dAB = a4 - a3;
                dir(0, sAB + E, dAB - E, sAB - E, dAB + E);
            }
            {   // j = 1: A = -(a0 + b1), B = a2, E = a4
                const double A = -(a0 + b1), sAB = A + a2, dAB = A - a2;
                dir(1, sAB + a4, dAB - a4, sAB - a4, dAB + a4);
            }
            {   // j = 2: A = -2 a0, B = a2, C = -a3
                const double m = -(2.0 * a0 + a3), n = a3 - 2.0 * a0;
                dir(2, m + a2, m - a2, n + a2, n - a2);
            }
            const double ka = pc->wk[0], kb = pc->wk[1];
            ed[0] = ka * fma(-2.0, Wv[2][0], Wv[0][3] - Wv[1][0]);
            ed[1] = -kb * (Wv[0][3] + Wv[1][0]);
            ed[2] = ka * (Wv[1][1] + Wv[2][1]);
            ed[3] = ka * (Wv[0][1] - Wv[2][2]);
            ed[4] = ka * (Wv[0][0] + Wv[1][3]);
            // D (upper triangle, row major) without its factors
            const double s01 = Wu[0][0] + Wu[1][0];
            D[0] = fma(4.0, Wu[2][0], s01);
            D[1] = Wu[1][0] - Wu[0][0];
            D[2] = -fma(2.0, Wu[2][1], Wu[1][1]);
            D[3] = fma(2.0, Wu[2][2], Wu[0][2]);
            D[4] = Wu[0][3] - Wu[1][3];
            D[5] = s01;
            D[6] = -Wu[1][1];
            D[7] = -Wu[0][2];
            D[8] = -(Wu[0][3] + Wu[1][3]);
            D[9] = Wu[1][0] + Wu[2][0];
            D[10] = -Wu[2][3];
            D[11] = Wu[1][2];
            D[12] = Wu[0][0] + Wu[2][0];
            D[13] = Wu[0][1];
            D[14] = s01;
            s = 12;
        }
        for (; s + 1 < nsys; s += 2) {
            RELOAD_BARRIER();   // Schmid vectors are re-read from the scalar cache
            double P0[5], P1[5];
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                P0[a] = pc->px[s][a];
                P1[a] = pc->px[s + 1][a];
            }
            const double f0 = lds_ic[s][threadIdx.x], f1 = lds_ic[s + 1][threadIdx.x];
            const double b0 = TWO ? lds_ic[TWO ? 12 + s : 0][threadIdx.x] : f0;
            const double b1 = TWO ? lds_ic[TWO ? 13 + s : 0][threadIdx.x] : f1;
            double r10, r20, r11, r21;
            rates(P0, f0, b0, r10, r20);
            rates(P1, f1, b1, r11, r21);
            accum(P0, r10, r20);
            accum(P1, r11, r21);
        }
        if (s < nsys) {
            RELOAD_BARRIER();
            double P0[5];
#pragma unroll
            for (int a = 0; a < 5; ++a) P0[a] = pc->px[s][a];
            const double f0 = lds_ic[s][threadIdx.x];
            const double b0 = TWO ? lds_ic[TWO ? 12 + s : 0][threadIdx.x] : f0;
            double r10, r20;
            rates(P0, f0, b0, r10, r20);
            accum(P0, r10, r20);
        }
        RELOAD_BARRIER();
        // Newton step x = J^-1 R (evolution.c:70-101) with
        // J = I + C0 (S_local + dt D) and R = sigma - lambda + C0 (eps_tot - eps),
        // eps_tot = eps_p_t + dt edot + S_local sigma.  Both carry the left
        // factor C0 = C066: J = C0 K with K = K_c + dt D, K_c = S0 + S_local,
        // and C0^-1 R = g = K_c sigma + w + dt edot.  K is symmetric positive
        // definite (two compliances plus the PSD flow-rule Jacobian), so
        // x = K^-1 g is an LDL^T solve without pivoting.  g is formed from
        // sigma afresh every step: a running update of K_c sigma would
        // carry the rounding of the large early steps (where dt edot and
        // dt D x nearly cancel) into the converged stress.
        // The hydrostatic component decouples (cubic elasticity: K_c has no
        // coupling beyond rounding, D has no hydrostatic row), so its Newton
        // step lands on h6 = -w6 / K_c66 and the rest is a 5x5 system in the
        // deviatoric block of K_c,x: one triangle of it (15 loads, issued
        // together) serves the residual and the factorisation.  (Loaded at
        // the top of the step, before the flow rule, it spills registers.)
        double K[15];
#pragma unroll
        for (int q = 0; q < 15; ++q) K[q] = kcv[q * 64];
        double g[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double t = lds_w[i][threadIdx.x] + dt * ed[i];
#pragma unroll
            for (int j = 0; j < 5; ++j) t += K[i >= j ? lt6(i, j) : lt6(j, i)] * sig[j];
            g[i] = t;
            glast[i] = t;
        }
        {
            int q = 0;
#pragma unroll
            for (int a = 0; a < 5; ++a)
#pragma unroll
                for (int b2 = a; b2 < 5; ++b2) {
                    if (FCC) {
                        // factors of the Walsh form: dt gam0 n / 12, / (4 sqrt 3), / 4
                        const double f = pc->wk[(q == 1 || (q >= 6 && q <= 8)) ? 3 : (q == 5 ? 4 : 2)];
                        K[lt6(b2, a)] += f * D[q];
                    } else {
                        K[lt6(b2, a)] += dt * D[q];
                    }
                    ++q;
                }
        }
        double x[5];
        if (ldlt_solve_packed<5>(K, g, x)) return 35;
        double cn = (h6 - sig[5]) * (h6 - sig[5]);
        sig[5] = h6;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const double so = sig[i];
            sig[i] -= x[i];
            cn += (sig[i] - so) * (sig[i] - so);
        }
        more = cn > tol2;
    }
    return 0;
}

template <int NT, bool UNI, bool TWO>
__global__ __launch_bounds__(NT, TWO ? 2 : VOCE_NR_WAVES) void k_nr(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    RVE_PROLOGUE(lb);
    MacroState *ms = r.ms;
    if (!iter_active(ms)) return;
    // per-lane values that stay constant through the Newton loop live in LDS
    // (component-major: conflict-free), keeping the register file for the loop
    __shared__ double lds_all[12 + (TWO ? 24 : 12)][NT];
    double (*lds_w)[NT] = lds_all;         // w = (eps_p - eps) - C066^-1 lambda
    double (*lds_old)[NT] = lds_all + 6;   // stress at the last flow-rule evaluation
    double (*lds_ic)[NT] = lds_all + 12;   // 1 / CRSS, forward (and backward)
    const long V = r.V;
    const long v = (long)lb * NT + threadIdx.x;
    double acc[14];
#pragma unroll
    for (int i = 0; i < 14; ++i) acc[i] = 0.0;
    if (v < V) {
        const uint32_t ori = as_global(r.ORI)[v];
        const int ph = as_global(r.PH)[v];
        if (ori == VOCE_GAS) {
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                as_global_mut(r.S6)[i * V + v] = 0.0;
                as_global_mut(r.P6)[i * V + v] = 0.0;
            }
        } else {
            const cptr<PhaseConst> pc = as_const(r.phase + (UNI ? 0 : (ph - 1)));
            const int nsys = pc->nsys;
            const cptr<MacroState> msc = as_const(ms);
            const double dt = msc->dt;
            const gptr<double> tab = as_global(r.SLOC + (long)ori * VOCE_OTAB);   // S_local, K_c, initial TranMat
            // Every load of the prologue goes out first, before any branch
            // or use: a wait on a load also waits on the loads issued before
            // it, and a load behind a branch or a use costs a round trip of
            // its own.  The CRSS arrays hold twelve systems per voxel.
            double cq[12], Rm[9], s0[6], w[6];
            const double kc66 = tab[36 + 35];   // K_c, hydrostatic diagonal entry
            {
                const gptr<double> cf = as_global(r.CRSS + v);
#pragma unroll
                for (int q = 0; q < 12; ++q) cq[q] = cf[(long)q * V];
#pragma unroll
                for (int i = 0; i < 9; ++i) Rm[i] = as_global(r.R)[i * V + v];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    s0[i] = as_global(r.S6)[i * V + v];
                    w[i] = (i < 5 ? as_global(r.EP5)[i * V + v] : 0.0) - as_global(r.E6)[i * V + v];
                }
            }
            // the flow rule only needs 1/CRSS (the same quotient 1.0/tau_c it
            // would form per evaluation); systems past nsys (CRSS 0) are
            // never read
#pragma unroll
            for (int q = 0; q < 12; ++q) lds_ic[q][threadIdx.x] = 1.0 / cq[q];
            if (TWO) {
                const gptr<double> cb = as_global(r.CRSSB + v);
#pragma unroll
                for (int q = 0; q < 12; ++q) cq[q] = cb[(long)q * V];
#pragma unroll
                for (int q = 0; q < 12; ++q) lds_ic[TWO ? 12 + q : q][threadIdx.x] = 1.0 / cq[q];
            }
            // The Newton loop runs in the crystal frame of TranMat R, which is
            // fixed during an increment: with Q the B-basis rotation of R and
            // Q6 = diag(Q, 1), sigma_x = Q6^T sigma, and every sample-frame
            // quantity of the step maps the same way (g_x = Q6^T g,
            // K_x = Q6^T K Q6 = K_c,x + dt D_x).  The Schmid vectors are then
            // the constant crystal-frame ones (Schm_xt, init.c:217-224) and
            // K_c,x is the per-voxel table the close-out keeps; the loop
            // never touches R.
            double sig[6], lam2 = 0.0;
            {
#pragma unroll
                for (int i = 0; i < 6; ++i) lam2 += s0[i] * s0[i];
                // w = (eps_p_t - eps) - S0 lambda (S0 = C066^-1): the constant
                // part of the compliance-form residual below, kept in LDS (a
                // per-lane register extension)
                const cptr<double> S0b = msc->S066;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
#pragma unroll
                    for (int j = 0; j < 6; ++j) w[i] -= S0b[i * 6 + j] * s0[j];
                }
                // to the crystal frame
                double wx[5];
                rot_dev<false>(Rm, s0, sig);
                rot_dev<false>(Rm, w, wx);
#pragma unroll
                for (int b = 0; b < 5; ++b) lds_w[b][threadIdx.x] = wx[b];
                sig[5] = s0[5];
                lds_w[5][threadIdx.x] = w[5];
            }
            const double tol2 = 1E-14 * lam2, h6 = -w[5] / kc66;
            if (msc->iter == 0) {
                // first iteration of the increment: K_c in the frame of this
                // increment's TranMat
                double Q[5][5];
                basis_rotation(Rm, Q);
                store_kc_crystal(r, v, r.SLOC + (long)ori * VOCE_OTAB + 36, Q);
            }
            const gptr<double> kcv = as_global(kcv_voxel(r, v));
            double ed[5], glast[5];
            // the power law's form is uniform per phase: one loop per form
            const int ni = (int)pc->nsr;
            const int pm = pc->int_n ? (ni == 20 ? PM_N20 : PM_INT) : PM_REAL;
            int err;
            if (pm == PM_N20 && pc->fcc12)
                err = newton_crystal<NT, TWO, PM_N20, true>(sig, ed, glast, tol2, h6, pc, nsys, dt, kcv, lds_w, lds_old, lds_ic);
            else if (pm == PM_N20)
                err = newton_crystal<NT, TWO, PM_N20, false>(sig, ed, glast, tol2, h6, pc, nsys, dt, kcv, lds_w, lds_old, lds_ic);
            else if (pm == PM_INT)
                err = newton_crystal<NT, TWO, PM_INT, false>(sig, ed, glast, tol2, h6, pc, nsys, dt, kcv, lds_w, lds_old, lds_ic);
            else
                err = newton_crystal<NT, TWO, PM_REAL, false>(sig, ed, glast, tol2, h6, pc, nsys, dt, kcv, lds_w, lds_old, lds_ic);
            if (err) ms->error = err;
            // Back to the sample frame: the stress, the stress of the last
            // flow-rule evaluation and the Newton residual g there.  The
            // error partial |eps_tot(sigma_old) - eps| (evolution.c:103-114)
            // is that of g(sigma_old) - C066^-1 (sigma_old - lambda), since
            // g = K_c sigma + w + dt edot, K_c = C066^-1 + S_local and
            // w = eps_p - eps - C066^-1 lambda.  The stress of the last
            // flow-rule evaluation is stored in the crystal frame (only the
            // close-out reads it).
            double osa[6], gsa[6];
            RELOAD_BARRIER();
            {
                double Rm[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) Rm[i] = as_global(r.R)[i * V + v];
                double sx[5], ox[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) ox[i] = lds_old[i][threadIdx.x];
#pragma unroll
                for (int i = 0; i < 5; ++i) sx[i] = sig[i];
#pragma unroll
                for (int i = 0; i < 6; ++i) as_global_mut(r.P6)[i * V + v] = ox[i];
                gsa[5] = fma(kc66, ox[5], lds_w[5][threadIdx.x]);
                osa[5] = ox[5];
                rot_dev<true>(Rm, sx, sig);
                rot_dev<true>(Rm, ox, osa);
                rot_dev<true>(Rm, glast, gsa);
            }
            double cs = 0.0, ce = 0.0;
            RELOAD_BARRIER();
            double lam[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                lam[i] = as_global(r.S6)[i * V + v];
                cs += (sig[i] - lam[i]) * (sig[i] - lam[i]);
                as_global_mut(r.S6)[i * V + v] = sig[i];
            }
            {
                const cptr<double> S0b = msc->S066;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    double t = gsa[i];
#pragma unroll
                    for (int j = 0; j < 6; ++j) t += S0b[i * 6 + j] * (lam[j] - osa[j]);
                    ce += t * t;
                }
            }
            acc[0] = sqrt(cs);
            acc[1] = sqrt(ce);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                acc[2 + i] = sig[i];
                acc[8 + i] = (ph == 1) ? sig[i] : 0.0;
            }
        }
    }
    if (UNI) {
        // every crystalline voxel is phase 1 (gas voxels add zeros to both)
        double a8[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) a8[i] = acc[i];
        block_reduce_store<NT, 8, true>(a8, lds_all, r.part + (long)lb * 16);
    } else {
        block_reduce_store<NT, 14>(acc, lds_all, r.part + (long)lb * 16);
    }
    if (lb == 0 && threadIdx.x == 0) ms->fin_pending = 1;   // read after this kernel ends
}

// Iteration finaliser: sums the Newton kernel's block partials in block
// order and runs the error norms, get_smacro and the convergence test, when
// a Newton step is pending.  It runs in one extra block of the next
// iteration's forward plane launch (whose other blocks only read S6 and write
// scratch spectra, so they may run beside it), or alone (k_fin_iter) after
// the last launched iteration.  The kernel boundary before it makes the
// Newton partials visible without a device-scope fence in every Newton
// block, which would write back and invalidate the L2 each time.
template <int NT>
__device__ __forceinline__ void finalize_block(const RveDev &r)
{
    MacroState *ms = r.ms;
    if (!ms->fin_pending) return;
    // the macro state comes into LDS with all threads while the partials
    // are summed, so the one-thread finaliser reads it at LDS latency
    __shared__ MacroState sms;
    static_assert(sizeof(MacroState) % 8 == 0, "MacroState is copied as doubles");
    {
        const gptr<double> src = as_global((const double *)ms);
        double *dst = (double *)&sms;
        grouped<NT, 4, double>((int)(sizeof(MacroState) / 8), [&](int i) { return src[i]; },
                               [&](int i, double v) { dst[i] = v; });
    }
    __shared__ double tot[16];
    __shared__ double dDs[9];
    sum_partials<NT, 14>(r.part, r.nblk_vox, r.uniform_phase != 0, tot);   // ends with a barrier
    if (threadIdx.x == 0) {
        finalize_iteration(ms, &sms, tot, dDs);
        ms->fin_pending = 0;
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_fin_iter(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    RVE_PROLOGUE(lb);
    (void)lb;
    finalize_block<NT>(r);
}

// increment start (evolution.c:306-340)
__global__ void k_begin_increment(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    RVE_PROLOGUE(lb);
    (void)lb;
    MacroState *ms = r.ms;
    if (ms->vg_pending) {
        // a new macroscopic velocity gradient for this increment: target
        // DisGradAvg_t + L dt (the predictor of evolution.c:460 with the new L)
        const double dt = ms->dt;
        for (int i = 0; i < 9; ++i) {
            const int t = (i % 3) * 3 + i / 3;
            ms->Udot[i] = ms->vg_next[i];
            ms->Udot_a[i] = (ms->vg_next[i] - ms->vg_next[t]) / 2.;
            ms->DisGradAvg[i] = ms->DisGradAvg_t[i] + ms->vg_next[i] * dt;
        }
        double sy[9];
        for (int i = 0; i < 9; ++i) sy[i] = (ms->vg_next[i] + ms->vg_next[(i % 3) * 3 + i / 3]) / 2.;
        sym_to_b6(sy, ms->udot_sym6);
        ms->vg_pending = 0;
    }
    ms->istep += 1;
    ms->iter = 0;
    ms->fin_pending = 0;
    for (int i = 0; i < 9; ++i) {
        ms->dD[i] = 0.0;
        ms->acum[i] = 0.0;
        ms->applied[i] = 0.0;
    }
    if (ms->creep) {
        double a[6];
        sym_to_b6(ms->DisGradAvg_actual, a);
        for (int i = 0; i < 6; ++i) {
            ms->creep_sym6[i] = a[i];
            ms->shift6[i] = 0.0;
        }
        const double *A = ms->DisGradAvg_actual;
        ms->creep_skew[0] = (A[7] - A[5]) / 2.;
        ms->creep_skew[1] = (A[2] - A[6]) / 2.;
        ms->creep_skew[2] = (A[3] - A[1]) / 2.;
    } else {
        double U[9];
        for (int i = 0; i < 9; ++i) U[i] = ms->Udot[i] * ms->dt;
        double a[6];
        sym_to_b6(U, a);
        for (int i = 0; i < 6; ++i) ms->shift6[i] = a[i];
    }
    // evolution.c:338-342: Err_e = Err_s = 2 Err before the first test
    ms->active = (ms->iter_max > 0) && (fabs(2. * ms->err_tol) > fabs(ms->err_tol));
    ms->error = 0;
}

// Orientation, constitutive.c:117-169
__device__ void orientation(double tran[9], const double rot_in[9])
{
    double rot[9];
    for (int i = 0; i < 9; ++i) rot[i] = rot_in[i];
    double w[3] = {rot[7], rot[2], rot[3]};
    double w_norm = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double w_tan = tan(w_norm / 2.0);
    if (fabs(w_norm) < 1E-6) w_norm = 1.0;
    for (int i = 0; i < 3; ++i) w[i] *= w_tan / w_norm;
    double om[9] = {0., -1. * w[2], w[1], w[2], 0., -1. * w[0], -1. * w[1], w[0], 0.};
    double om2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            om2[i * 3 + j] = om[i * 3 + 0] * om[0 * 3 + j] + om[i * 3 + 1] * om[1 * 3 + j] + om[i * 3 + 2] * om[2 * 3 + j];
    w_tan = w_tan * w_tan;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) rot[i * 3 + j] = (double)(i == j) + 2. * (om[i * 3 + j] + om2[i * 3 + j]) / (1. + w_tan);
    double tn[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            tn[i * 3 + j] = rot[i * 3 + 0] * tran[0 * 3 + j] + rot[i * 3 + 1] * tran[1 * 3 + j] + rot[i * 3 + 2] * tran[2 * 3 + j];
    for (int i = 0; i < 9; ++i) tran[i] = tn[i];
}

// ------------------------------------------------------------------
// Increment close-out per voxel (evolution.c:433-530).  Partials per
// block: [0..4] sum Eps (B basis), [5..9] sum Edot.
// ------------------------------------------------------------------
template <int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_closeout(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    RVE_PROLOGUE(lb);
    const MacroState *ms = r.ms;
    const bool rot_on = rve_need_skew(ms);
    const long V = r.V;
    const long v = (long)lb * NT + threadIdx.x;
    double acc[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) acc[i] = 0.0;
    const int istep = ms->istep;
    const double dt = ms->dt;
    const int diag = ms->diag;
    if (v < V) {
        const uint32_t ori = r.ORI[v];
        const int ph = r.PH[v];
        // spin of the increment: skew(DisGrad_new - DisGrad_old), components (2,1),(0,2),(1,0)
        double dsk[3] = {0.0, 0.0, 0.0};
        if (rot_on) {
            const double *A = ms->applied;
            const double sa[3] = {(A[7] - A[5]) / 2., (A[2] - A[6]) / 2., (A[3] - A[1]) / 2.};
            double base[3];
            if (ms->creep) {
#pragma unroll
                for (int i = 0; i < 3; ++i) base[i] = ms->creep_skew[i] - r.WG[i * V + v];
            } else {
                const double *U = ms->Udot;
                base[0] = (U[7] - U[5]) / 2. * dt;
                base[1] = (U[2] - U[6]) / 2. * dt;
                base[2] = (U[3] - U[1]) / 2. * dt;
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                dsk[i] = base[i] + sa[i] + r.WF[i * V + v];
                r.WG[i * V + v] += dsk[i];
                r.WF[i * V + v] = dsk[i];
            }
        }
        if (ori == VOCE_GAS) {
#pragma unroll
            for (int i = 0; i < 5; ++i) r.ED5[i * V + v] = 0.0;
#pragma unroll
            for (int i = 0; i < 5; ++i) acc[i] = r.EP5[i * V + v];
        } else {
            const PhaseConst *pc = r.phase + (UNI ? 0 : (ph - 1));
            const int nsys = pc->nsys;
            double Rm[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) Rm[i] = r.R[i * V + v];
            double Q[5][5];
            basis_rotation(Rm, Q);
            // the stress of the last flow-rule evaluation, crystal frame
            double sx6[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) sx6[i] = r.P6[i * V + v];
            double *crf = r.CRSS;
            double *crb = pc->two_crss ? r.CRSSB : r.CRSS;
            double cf[12], cb[12];
#pragma unroll
            for (int s = 0; s < 12; ++s) {
                cf[s] = s < nsys ? crf[s * V + v] : 1.0;
                cb[s] = s < nsys ? crb[s * V + v] : 1.0;
            }
            double ex[5], ed[5], gd[12], tx[12];
#pragma unroll
            for (int s = 0; s < 12; ++s) gd[s] = tx[s] = 0.0;
            flow_rule_x(sx6, pc, nsys, cf, cb, ex, gd, tx);
            // strain rate and the pre-update stress in the sample frame
            double sp[6];
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                ed[a] = Q[a][0] * ex[0] + Q[a][1] * ex[1] + Q[a][2] * ex[2] + Q[a][3] * ex[3] + Q[a][4] * ex[4];
                sp[a] = Q[a][0] * sx6[0] + Q[a][1] * sx6[1] + Q[a][2] * sx6[2] + Q[a][3] * sx6[3] + Q[a][4] * sx6[4];
            }
            sp[5] = sx6[5];
            double ep[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) ep[i] = r.EP5[i * V + v];
            if (diag) {
                // EVM from eps_tot at the pre-update stress, SVM from sigma (evolution.c:134-155)
                const double *Sl = r.SLOC + (long)ori * VOCE_OTAB;
                double tt[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    double t = (i < 5 ? ep[i] + ed[i] * dt : 0.0);
#pragma unroll
                    for (int j = 0; j < 6; ++j) t += Sl[i * 6 + j] * sp[j];
                    tt[i] = t;
                }
                double T[9];
                b6_to_t(tt, T);
                r.EVM[v] = von_mises_strain(T);
                double s6[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) s6[i] = r.S6[i * V + v];
                double ST[9];
                b6_to_t(s6, ST);
                const double tr = ST[0] + ST[4] + ST[8];
                double sv = 0.0;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const double d = (ST[i * 3 + j] + ST[j * 3 + i]) / 2. - (double)(i == j) * tr / 3.0;
                        sv += d * d;
                    }
                r.SVM[v] = sqrt(3. / 2. * sv);
#pragma unroll
                for (int s = 0; s < 12; ++s) {
                    r.GAMDOT[s * V + v] = gd[s];
                    r.TAUX[s * V + v] = tx[s];
                }
            }
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                ep[i] += ed[i] * dt;
                r.EP5[i * V + v] = ep[i];
                r.ED5[i * V + v] = ed[i];
                acc[i] = ep[i];
                acc[5 + i] = ed[i];
            }
            // update_orient (constitutive.c:171-248), from step 2
            if (ms->update_flag == 1 && istep > 1) {
                double Lp[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) Lp[i] = 0.0;
#pragma unroll
                for (int s = 0; s < 12; ++s) {
                    if (s < nsys) {
                        double bsa[3], nsa[3];
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            bsa[i] = 0.0;
                            nsa[i] = 0.0;
#pragma unroll
                            for (int j = 0; j < 3; ++j) {
                                bsa[i] += Rm[i * 3 + j] * pc->rb[s][j];
                                nsa[i] += Rm[i * 3 + j] * pc->rn[s][j];
                            }
                        }
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int j = 0; j < 3; ++j) Lp[i * 3 + j] += bsa[i] * nsa[j] * gd[s];
                    }
                }
                const double RotLoc[9] = {0.0, -dsk[2] / dt, dsk[1] / dt, dsk[2] / dt, 0.0, -dsk[0] / dt,
                                          -dsk[1] / dt, dsk[0] / dt, 0.0};
                double Rot[9];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const double rs = (Lp[i * 3 + j] - Lp[j * 3 + i]) / 2.;
                        Rot[i * 3 + j] = (ms->Udot_a[i * 3 + j] + RotLoc[i * 3 + j] - rs) * dt;
                    }
                double Rn[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) Rn[i] = Rm[i];
                orientation(Rn, Rot);
#pragma unroll
                for (int i = 0; i < 9; ++i) r.R[i * V + v] = Rn[i];
            }
            // harden (constitutive.c:250-310), from step 3, and the
            // rho/work diagnostics that follow it (evolution.c:509-529)
            if (ms->hard_flag == 1 && istep > 2) {
                const double gamtot = r.GAC[v];
                double deltgam = 0.0;
#pragma unroll
                for (int s = 0; s < 12; ++s)
                    if (s < nsys) r.GTOT[s * V + v] = r.GTOT[s * V + v] + gd[s] * dt;
#pragma unroll
                for (int s = 0; s < 12; ++s)
                    if (s < nsys) deltgam += fabs(gd[s]) * dt;
                const double tau0 = pc->tau0f, tau1 = pc->tau1, thet0 = pc->thet0, thet1 = pc->thet1;
                const double TINY = 1.E-4 * tau0;
                double voce = 0.0;
                if (fabs(thet0) > TINY) {
                    voce = thet1 * deltgam;
                    if (fabs(tau1) > TINY) {
                        const double fact = fabs(thet0 / tau1);
                        const double exp_ini = exp(-1.0 * gamtot * fact);
                        const double exp_del = exp(-1.0 * deltgam * fact);
                        voce += -1. * (fact * tau1 - thet1) / fact * exp_ini * (exp_del - 1.0) -
                                thet1 / fact * exp_ini * (exp_del * ((gamtot + deltgam) * fact + 1.0) - (gamtot * fact + 1.0));
                    }
                }
#pragma unroll
                for (int s = 0; s < 12; ++s) {
                    if (s < nsys) {
                        double dtau = 0.;
#pragma unroll
                        for (int j = 0; j < 12; ++j)
                            if (j < nsys) dtau += pc->hard[s][j] * fabs(gd[j]) * dt;
                        crf[s * V + v] += dtau * voce / deltgam;
                        if (pc->two_crss) crb[s * V + v] += dtau * voce / deltgam;
                    }
                }
                r.GAC[v] = gamtot + deltgam;
                if (diag) {
                    double s6[6];
#pragma unroll
                    for (int i = 0; i < 6; ++i) s6[i] = r.S6[i * V + v];
                    double sx[5];
#pragma unroll
                    for (int b = 0; b < 5; ++b)
                        sx[b] = Q[0][b] * s6[0] + Q[1][b] * s6[1] + Q[2][b] * s6[2] + Q[3][b] * s6[3] + Q[4][b] * s6[4];
                    double rr = 0.0, rc = 0.0, wp = 0.0;
                    const double K = (0.35 * 47.565 * 0.256);
#pragma unroll
                    for (int s = 0; s < 12; ++s) {
                        if (s < nsys) {
                            const double *p = pc->px[s];
                            const double rss = p[0] * sx[0] + p[1] * sx[1] + p[2] * sx[2] + p[3] * sx[3] + p[4] * sx[4];
                            rr += (rss / K) * (rss / K);
                            rc += (crf[s * V + v] / K) * (crf[s * V + v] / K);
                            wp += fabs((r.GTOT[s * V + v]) * rss);
                        }
                    }
                    r.RHOR[v] = rr;
                    r.RHOC[v] = rc;
                    r.WORKP[v] = wp;
                }
            }
        }
    }
    __shared__ double red[10][NT];
    block_reduce_store<NT, 10>(acc, red, r.part + (long)lb * 16);
}

// increment finaliser (evolution.c:438-498)
template <int NT>
__global__ __launch_bounds__(NT) void k_fin_incr(const RveDev *__restrict__ rtab, const int2 *__restrict__ bmap)
{
    RVE_PROLOGUE(lb);
    (void)lb;
    MacroState *ms = r.ms;
    __shared__ double tot[16];
    sum_partials<NT, 10>(r.part, r.nblk_vox, false, tot);
    if (threadIdx.x != 0) return;
    const double dt = ms->dt;
    for (int i = 0; i < 9; ++i) {
        ms->DisGradAvg_actual[i] = ms->DisGradAvg[i] + ms->acum[i];
        ms->VelGradAvg[i] = (ms->DisGradAvg_actual[i] - ms->DisGradAvg_t[i]) / dt;
    }
    ms->e_vm = von_mises_strain(ms->DisGradAvg_actual);
    ms->d_vm = von_mises_strain(ms->VelGradAvg);
    for (int i = 0; i < 9; ++i) {
        ms->DisGradAvg_t[i] = ms->DisGradAvg_actual[i];
        ms->DisGradAvg[i] = ms->DisGradAvg_t[i] + ms->Udot[i] * dt;
    }
    double e6[6] = {tot[0], tot[1], tot[2], tot[3], tot[4], 0.0};
    double d6[6] = {tot[5], tot[6], tot[7], tot[8], tot[9], 0.0};
    double ET[9], DT[9];
    b6_to_t(e6, ET);
    b6_to_t(d6, DT);
    double evmp = 0.0, dvmp = 0.0;
    for (int i = 0; i < 9; ++i) {
        ms->EpsAvg[i] = ET[i] * ms->wgt;
        ms->EdotAvg[i] = DT[i] * ms->wgt;
        evmp += ms->EpsAvg[i] * ms->EpsAvg[i];
        dvmp += ms->EdotAvg[i] * ms->EdotAvg[i];
    }
    ms->evmp = sqrt(2. / 3. * evmp);
    ms->dvmp = sqrt(2. / 3. * dvmp);
}

// Batch convergence poll: out[0] = members still iterating, out[1] = the
// largest error code among them (0: none).
__global__ __launch_bounds__(256) void k_count_active(const RveDev *__restrict__ tab, int n, int *out)
{
    __shared__ int s_act, s_err;
    if (threadIdx.x == 0) {
        s_act = 0;
        s_err = 0;
    }
    __syncthreads();
    int a = 0, e = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const MacroState *ms = tab[i].ms;
        a += ms->active != 0;
        e = max(e, ms->error);
    }
    atomicAdd(&s_act, a);
    atomicMax(&s_err, e);
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = s_act;
        out[1] = s_err;
    }
}

// Initial state (InitFields, init.c:486-534; CRSS init.c:789-792): from the
// orientation index and phase of each voxel, the TranMat and stress of its
// orientation-table row, CRSS = tau0 on the phase's systems, zeros elsewhere.
__global__ void k_init_state(const RveDev *__restrict__ rtab, const double *sig6_by_orient)
{
    const RveDev &r = rtab[0];
    const long V = r.V;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t ori = r.ORI[v];
    const bool gas = (ori == VOCE_GAS);
    const PhaseConst *pc = r.phase + (r.PH[v] - 1);
    const double *row = r.SLOC + (long)(gas ? 0 : ori) * VOCE_OTAB + 72;
    for (int i = 0; i < 9; ++i) r.R[i * V + v] = gas ? (double)(i % 4 == 0) : row[i];
    for (int i = 0; i < 6; ++i) {
        r.S6[i * V + v] = gas ? 0.0 : sig6_by_orient[(long)ori * 6 + i];
        r.E6[i * V + v] = 0.0;
        if (r.P6) r.P6[i * V + v] = 0.0;
    }
    for (int i = 0; i < 5; ++i) {
        r.EP5[i * V + v] = 0.0;
        r.ED5[i * V + v] = 0.0;
    }
    for (int i = 0; i < 3; ++i) {
        r.WG[i * V + v] = 0.0;
        r.WF[i * V + v] = 0.0;
    }
    r.GAC[v] = 0.0;
    const int nsys = gas ? 0 : pc->nsys;
    for (int s = 0; s < 12; ++s) {
        r.GTOT[s * V + v] = 0.0;
        r.CRSS[s * V + v] = (s < nsys) ? pc->tau0f : 0.0;
        if (r.CRSSB != r.CRSS) r.CRSSB[s * V + v] = (s < nsys) ? pc->tau0b : 0.0;
    }
}

#ifndef VOCE_NT_VOX
#define VOCE_NT_VOX 256
#endif
constexpr int NT_VOX = VOCE_NT_VOX;   // threads per block of the per-voxel kernels
constexpr int NT_FIN = 256;
// Launch-shape variants (threads per block, largest tile in complex
// elements).  Plane kernels: ny * (nz/2+1) <= MAXEL.  x-pass kernel:
// 6 * tk * nx <= MAXEL.  Which one runs for a grid size is tuned.
constexpr int NT_PL_T = 64, MAXEL_PL_T = 576;     // planes up to 32 x 32
constexpr int NT_PL_S = 256, MAXEL_PL_S = 2112;   // planes up to 64 x 64
constexpr int NT_PL_L = 512, MAXEL_PL_L = 8320;   // planes up to 128 x 128
// 64 x 64 with five waves: the 33 half-spectrum columns of a 64 x 64 plane
// are 264 radix-8 butterflies per stage, one per thread
constexpr int NT_PL_M = 320, MAXEL_PL_M = 2112;
constexpr int NT_X = 256, MAXEL_X = 3072;
constexpr int NT_XB = 128, MAXEL_XB = 1536;
constexpr int NT_XC = 64, MAXEL_XC = 768;

// Events the next iteration-kernel launch binds to its own dispatch
// (hipExtLaunchKernelGGL): the first takes the kernel's start time, the
// second its end, so their elapsed time is the kernel's execution without
// the queue's launch gaps.  Null outside timed launches, which then go
// through the plain launch (graph capture included).
thread_local hipEvent_t g_kt0 = nullptr, g_kt1 = nullptr;
// chunks of the split passes for the set's planes: row pairs of the z passes
// (a pair is 2 (Nz/2 + 1) elements), kz columns of the y pass (Ny elements each)
static inline unsigned sp_chunks_z(const LaunchSet *L)
{
    const int np = (L->sp_ny + 1) / 2, ppc = MAXEL_SP / (2 * (L->sp_nz / 2 + 1));
    return (unsigned)((np + ppc - 1) / ppc);
}
static inline unsigned sp_chunks_y(const LaunchSet *L)
{
    const int nzh = L->sp_nz / 2 + 1, kpc = MAXEL_SP / L->sp_ny;
    return (unsigned)((nzh + kpc - 1) / kpc);
}

#define VLAUNCH(kern, grid, block, shm, st, ...)                                           \
    do {                                                                                    \
        if (g_kt0) hipExtLaunchKernelGGL(kern, grid, block, shm, st, g_kt0, g_kt1, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(kern, grid, block, shm, st, __VA_ARGS__);                   \
    } while (0)

template <typename T>
void launch_plane_fwd(const LaunchSet *L, hipStream_t st)
{
    switch (L->plv) {
    case 0: VLAUNCH((k_plane_fwd<T, NT_PL_T, MAXEL_PL_T>), dim3(L->n_pfwd), dim3(NT_PL_T), L->pl_lds, st, L->tab, L->map_pfwd); break;
    case 1: VLAUNCH((k_plane_fwd<T, NT_PL_S, MAXEL_PL_S>), dim3(L->n_pfwd), dim3(NT_PL_S), L->pl_lds, st, L->tab, L->map_pfwd); break;
    case 3: VLAUNCH((k_plane_fwd<T, NT_PL_M, MAXEL_PL_M>), dim3(L->n_pfwd), dim3(NT_PL_M), L->pl_lds, st, L->tab, L->map_pfwd); break;
    case 4: {   // line-parallel, square power-of-two planes (the grid's ny)
        const int n = L->p2_n;
        if (n == 16) VLAUNCH((k_plane_fwd_p2<T, 16>), dim3(L->n_pfwd), dim3(P2Plan<16>::NT), L->pl_lds, st, L->tab, L->map_pfwd);
        else if (n == 32) VLAUNCH((k_plane_fwd_p2<T, 32>), dim3(L->n_pfwd), dim3(P2Plan<32>::NT), L->pl_lds, st, L->tab, L->map_pfwd);
        else if (n == 48) VLAUNCH((k_plane_fwd_p2<T, 48>), dim3(L->n_pfwd), dim3(P2Plan<48>::NT), L->pl_lds, st, L->tab, L->map_pfwd);
        else if (n == 64) VLAUNCH((k_plane_fwd_p2<T, 64>), dim3(L->n_pfwd), dim3(P2Plan<64>::NT), L->pl_lds, st, L->tab, L->map_pfwd);
        else VLAUNCH((k_plane_fwd_p2<T, 128>), dim3(L->n_pfwd), dim3(P2Plan<128>::NT), L->pl_lds, st, L->tab, L->map_pfwd);
        break;
    }
    case 5: {   // split z and y passes (planes too large for one block), a block per chunk; a
                // timed pass binds the start event to the first launch and the end event to the second
        const dim3 gz(L->n_pfwd, sp_chunks_z(L)), gy(L->n_plane6, sp_chunks_y(L));
        if (g_kt0) {
            hipExtLaunchKernelGGL((k_zfwd_s<T>), gz, dim3(NT_SP), L->pl_lds, st, g_kt0, nullptr, 0, L->tab, L->map_pfwd);
            hipExtLaunchKernelGGL((k_y_s<T, false>), gy, dim3(NT_SP), L->pl_lds, st, nullptr, g_kt1, 0, L->tab,
                                  L->map_plane6, 0);
        } else {
            hipLaunchKernelGGL((k_zfwd_s<T>), gz, dim3(NT_SP), L->pl_lds, st, L->tab, L->map_pfwd);
            hipLaunchKernelGGL((k_y_s<T, false>), gy, dim3(NT_SP), L->pl_lds, st, L->tab, L->map_plane6, 0);
        }
        break;
    }
    case 6: {   // line-parallel split passes (256 x 256 planes)
        const dim3 gz(L->n_pfwd, SL_N / 2 / SL_LPB), gy(L->n_plane6, (SL_N / 2 + SLY<T>::LPB) / SLY<T>::LPB);
        const size_t zl = sl_lds_z<T>(), yl = sl_lds_y<T>();
        if (g_kt0) {
            hipExtLaunchKernelGGL((k_zfwd_l<T>), gz, dim3(SL_NT), zl, st, g_kt0, nullptr, 0, L->tab, L->map_pfwd);
            hipExtLaunchKernelGGL((k_y_l<T, false>), gy, dim3(SLY<T>::NT), yl, st, nullptr, g_kt1, 0, L->tab,
                                  L->map_plane6, 0);
        } else {
            hipLaunchKernelGGL((k_zfwd_l<T>), gz, dim3(SL_NT), zl, st, L->tab, L->map_pfwd);
            hipLaunchKernelGGL((k_y_l<T, false>), gy, dim3(SLY<T>::NT), yl, st, L->tab, L->map_plane6, 0);
        }
        break;
    }
    default: VLAUNCH((k_plane_fwd<T, NT_PL_L, MAXEL_PL_L>), dim3(L->n_pfwd), dim3(NT_PL_L), L->pl_lds, st, L->tab, L->map_pfwd); break;
    }
}

template <typename T>
void launch_plane_inv(const LaunchSet *L, hipStream_t st, int mode)
{
    const int n = mode ? L->n_plane3 : L->n_plane6;
    const int2 *map = mode ? L->map_plane3 : L->map_plane6;
    switch (L->plv) {
    case 0: VLAUNCH((k_plane_inv<T, NT_PL_T, MAXEL_PL_T>), dim3(n), dim3(NT_PL_T), L->pl_lds, st, L->tab, map, mode); break;
    case 1: VLAUNCH((k_plane_inv<T, NT_PL_S, MAXEL_PL_S>), dim3(n), dim3(NT_PL_S), L->pl_lds, st, L->tab, map, mode); break;
    case 3: VLAUNCH((k_plane_inv<T, NT_PL_M, MAXEL_PL_M>), dim3(n), dim3(NT_PL_M), L->pl_lds, st, L->tab, map, mode); break;
    case 4: {
        const int m = L->p2_n;
        if (m == 16) VLAUNCH((k_plane_inv_p2<T, 16>), dim3(n), dim3(P2Plan<16>::NT), L->pl_lds, st, L->tab, map, mode);
        else if (m == 32) VLAUNCH((k_plane_inv_p2<T, 32>), dim3(n), dim3(P2Plan<32>::NT), L->pl_lds, st, L->tab, map, mode);
        else if (m == 48) VLAUNCH((k_plane_inv_p2<T, 48>), dim3(n), dim3(P2Plan<48>::NT), L->pl_lds, st, L->tab, map, mode);
        else if (m == 64) VLAUNCH((k_plane_inv_p2<T, 64>), dim3(n), dim3(P2Plan<64>::NT), L->pl_lds, st, L->tab, map, mode);
        else VLAUNCH((k_plane_inv_p2<T, 128>), dim3(n), dim3(P2Plan<128>::NT), L->pl_lds, st, L->tab, map, mode);
        break;
    }
    case 5: {   // split y and z passes, a block per chunk
        const dim3 gy(n, sp_chunks_y(L)), gz(n, sp_chunks_z(L));
        if (g_kt0) {
            hipExtLaunchKernelGGL((k_y_s<T, true>), gy, dim3(NT_SP), L->pl_lds, st, g_kt0, nullptr, 0, L->tab, map, mode);
            hipExtLaunchKernelGGL((k_zinv_s<T>), gz, dim3(NT_SP), L->pl_lds, st, nullptr, g_kt1, 0, L->tab, map, mode);
        } else {
            hipLaunchKernelGGL((k_y_s<T, true>), gy, dim3(NT_SP), L->pl_lds, st, L->tab, map, mode);
            hipLaunchKernelGGL((k_zinv_s<T>), gz, dim3(NT_SP), L->pl_lds, st, L->tab, map, mode);
        }
        break;
    }
    case 6: {   // line-parallel split passes (256 x 256 planes)
        const dim3 gy(n, (SL_N / 2 + SLY<T>::LPB) / SLY<T>::LPB), gz(n, SL_N / 2 / SL_LPB);
        const size_t zl = sl_lds_z<T>(), yl = sl_lds_y<T>();
        if (g_kt0) {
            hipExtLaunchKernelGGL((k_y_l<T, true>), gy, dim3(SLY<T>::NT), yl, st, g_kt0, nullptr, 0, L->tab, map, mode);
            hipExtLaunchKernelGGL((k_zinv_l<T>), gz, dim3(SL_NT), zl, st, nullptr, g_kt1, 0, L->tab, map, mode);
        } else {
            hipLaunchKernelGGL((k_y_l<T, true>), gy, dim3(SLY<T>::NT), yl, st, L->tab, map, mode);
            hipLaunchKernelGGL((k_zinv_l<T>), gz, dim3(SL_NT), zl, st, L->tab, map, mode);
        }
        break;
    }
    default: VLAUNCH((k_plane_inv<T, NT_PL_L, MAXEL_PL_L>), dim3(n), dim3(NT_PL_L), L->pl_lds, st, L->tab, map, mode); break;
    }
}

template <typename T, int NX>
void launch_xgreen_p2(const LaunchSet *L, hipStream_t st, int mode)
{
    const int wpb = L->p2_tk / XP2<NX>::CPW;
    if (wpb == 4) VLAUNCH((k_xgreen_p2<T, NX, 4>), dim3(L->n_x), dim3(256), L->x_lds, st, L->tab, L->map_x, mode);
    else VLAUNCH((k_xgreen_p2<T, NX, 2>), dim3(L->n_x), dim3(128), L->x_lds, st, L->tab, L->map_x, mode);
}

template <typename T, int NX>
void launch_xgreen_c(const LaunchSet *L, hipStream_t st, int mode)
{
    constexpr int TB = 128 / (int)sizeof(typename cplx_of<T>::t);
    if (L->p2_tk == TB)
        VLAUNCH((k_xgreen_c<T, NX, TB>), dim3(L->n_x), dim3(XC<T, NX, TB>::NT), L->x_lds, st, L->tab, L->map_x, mode);
    else if (L->p2_tk == TB / 2)
        VLAUNCH((k_xgreen_c<T, NX, TB / 2>), dim3(L->n_x), dim3(XC<T, NX, TB / 2>::NT), L->x_lds, st, L->tab, L->map_x,
                mode);
    else
        VLAUNCH((k_xgreen_c<T, NX, TB / 4>), dim3(L->n_x), dim3(XC<T, NX, TB / 4>::NT), L->x_lds, st, L->tab, L->map_x,
                mode);
}

// Nx = 256: half or quarter column blocks only (a whole one takes 1536
// threads in FP32, 196 KB of LDS in FP64)
template <typename T>
void launch_xgreen_c256(const LaunchSet *L, hipStream_t st, int mode)
{
    constexpr int TB = 128 / (int)sizeof(typename cplx_of<T>::t);
    if (L->p2_tk == TB / 2)
        VLAUNCH((k_xgreen_c<T, 256, TB / 2>), dim3(L->n_x), dim3(XC<T, 256, TB / 2>::NT), L->x_lds, st, L->tab, L->map_x,
                mode);
    else
        VLAUNCH((k_xgreen_c<T, 256, TB / 4>), dim3(L->n_x), dim3(XC<T, 256, TB / 4>::NT), L->x_lds, st, L->tab, L->map_x,
                mode);
}

template <typename T>
void launch_xgreen(const LaunchSet *L, hipStream_t st, int mode)
{
    switch (L->xv) {
    case 3:   // line-parallel x pass (the grid's nx)
        if (L->p2_nx == 16) launch_xgreen_p2<T, 16>(L, st, mode);
        else if (L->p2_nx == 32) launch_xgreen_p2<T, 32>(L, st, mode);
        else launch_xgreen_p2<T, 64>(L, st, mode);
        break;
    case 4:   // x pass staged through LDS (the grid's nx)
        if (L->p2_nx == 16) launch_xgreen_c<T, 16>(L, st, mode);
        else if (L->p2_nx == 32) launch_xgreen_c<T, 32>(L, st, mode);
        else if (L->p2_nx == 48) launch_xgreen_c<T, 48>(L, st, mode);
        else if (L->p2_nx == 64) launch_xgreen_c<T, 64>(L, st, mode);
        else if (L->p2_nx == 128) launch_xgreen_c<T, 128>(L, st, mode);
        else launch_xgreen_c256<T>(L, st, mode);
        break;
    case 0: VLAUNCH((k_xgreen<T, NT_X, MAXEL_X>), dim3(L->n_x), dim3(NT_X), L->x_lds, st, L->tab, L->map_x, mode); break;
    case 1: VLAUNCH((k_xgreen<T, NT_XB, MAXEL_XB>), dim3(L->n_x), dim3(NT_XB), L->x_lds, st, L->tab, L->map_x, mode); break;
    default: VLAUNCH((k_xgreen<T, NT_XC, MAXEL_XC>), dim3(L->n_x), dim3(NT_XC), L->x_lds, st, L->tab, L->map_x, mode); break;
    }
}

void launch_nr(const LaunchSet *L, hipStream_t st)
{
    if (L->uniform_phase) {
        if (L->two_crss) VLAUNCH((k_nr<NT_VOX, true, true>), dim3(L->n_vox), dim3(NT_VOX), 0, st, L->tab, L->map_vox);
        else VLAUNCH((k_nr<NT_VOX, true, false>), dim3(L->n_vox), dim3(NT_VOX), 0, st, L->tab, L->map_vox);
    } else {
        if (L->two_crss) VLAUNCH((k_nr<NT_VOX, false, true>), dim3(L->n_vox), dim3(NT_VOX), 0, st, L->tab, L->map_vox);
        else VLAUNCH((k_nr<NT_VOX, false, false>), dim3(L->n_vox), dim3(NT_VOX), 0, st, L->tab, L->map_vox);
    }
}

template <typename T>
hipError_t launch_iter_t(const LaunchSet *L, hipStream_t st, hipEvent_t *ev)
{
    // ev: null, or 8 events, a start/end pair per kernel in launch order
    auto bind = [&](int k) {
        g_kt0 = ev ? ev[2 * k] : nullptr;
        g_kt1 = ev ? ev[2 * k + 1] : nullptr;
    };
    bind(0);
    launch_plane_fwd<T>(L, st);
    bind(1);
    launch_xgreen<T>(L, st, 0);
    bind(2);
    launch_plane_inv<T>(L, st, 0);
    bind(3);
    launch_nr(L, st);
    g_kt0 = g_kt1 = nullptr;
    return hipGetLastError();
}

template <typename T>
hipError_t launch_skew_t(const LaunchSet *L, hipStream_t st)
{
    launch_xgreen<T>(L, st, 1);
    launch_plane_inv<T>(L, st, 1);
    return hipGetLastError();
}

template <typename T>
hipError_t set_lds_limits()
{
    const int big = 160 * 1024;
    const void *fns[] = {(const void *)k_plane_fwd<T, NT_PL_T, MAXEL_PL_T>, (const void *)k_plane_inv<T, NT_PL_T, MAXEL_PL_T>,
                         (const void *)k_plane_fwd<T, NT_PL_S, MAXEL_PL_S>, (const void *)k_plane_inv<T, NT_PL_S, MAXEL_PL_S>,
                         (const void *)k_plane_fwd<T, NT_PL_L, MAXEL_PL_L>, (const void *)k_plane_inv<T, NT_PL_L, MAXEL_PL_L>,
                         (const void *)k_plane_fwd<T, NT_PL_M, MAXEL_PL_M>, (const void *)k_plane_inv<T, NT_PL_M, MAXEL_PL_M>,
                         (const void *)k_xgreen<T, NT_X, MAXEL_X>, (const void *)k_xgreen<T, NT_XB, MAXEL_XB>,
                         (const void *)k_xgreen<T, NT_XC, MAXEL_XC>,
                         (const void *)k_plane_fwd_p2<T, 16>, (const void *)k_plane_inv_p2<T, 16>,
                         (const void *)k_plane_fwd_p2<T, 32>, (const void *)k_plane_inv_p2<T, 32>,
                         (const void *)k_plane_fwd_p2<T, 48>, (const void *)k_plane_inv_p2<T, 48>,
                         (const void *)k_plane_fwd_p2<T, 64>, (const void *)k_plane_inv_p2<T, 64>,
                         (const void *)k_plane_fwd_p2<T, 128>, (const void *)k_plane_inv_p2<T, 128>,
                         (const void *)k_xgreen_p2<T, 16, 2>, (const void *)k_xgreen_p2<T, 16, 4>,
                         (const void *)k_xgreen_p2<T, 32, 2>, (const void *)k_xgreen_p2<T, 32, 4>,
                         (const void *)k_xgreen_p2<T, 64, 2>, (const void *)k_xgreen_p2<T, 64, 4>,
                         (const void *)k_xgreen_c<T, 16, XC<T, 16, 8>::TB>, (const void *)k_xgreen_c<T, 16, XC<T, 16, 8>::TB / 2>,
                         (const void *)k_xgreen_c<T, 16, XC<T, 16, 8>::TB / 4>,
                         (const void *)k_xgreen_c<T, 32, XC<T, 32, 8>::TB>, (const void *)k_xgreen_c<T, 32, XC<T, 32, 8>::TB / 2>,
                         (const void *)k_xgreen_c<T, 32, XC<T, 32, 8>::TB / 4>,
                         (const void *)k_xgreen_c<T, 64, XC<T, 64, 8>::TB>, (const void *)k_xgreen_c<T, 64, XC<T, 64, 8>::TB / 2>,
                         (const void *)k_xgreen_c<T, 64, XC<T, 64, 8>::TB / 4>,
                         (const void *)k_xgreen_c<T, 48, XC<T, 48, 8>::TB>, (const void *)k_xgreen_c<T, 48, XC<T, 48, 8>::TB / 2>,
                         (const void *)k_xgreen_c<T, 48, XC<T, 48, 8>::TB / 4>,
                         (const void *)k_xgreen_c<T, 128, XC<T, 128, 8>::TB>, (const void *)k_xgreen_c<T, 128, XC<T, 128, 8>::TB / 2>,
                         (const void *)k_xgreen_c<T, 128, XC<T, 128, 8>::TB / 4>,
                         (const void *)k_xgreen_c<T, 256, XC<T, 256, 8>::TB / 2>, (const void *)k_xgreen_c<T, 256, XC<T, 256, 8>::TB / 4>,
                         (const void *)k_zfwd_s<T>, (const void *)k_y_s<T, false>, (const void *)k_y_s<T, true>,
                         (const void *)k_zinv_s<T>, (const void *)k_zfwd_l<T>, (const void *)k_y_l<T, false>,
                         (const void *)k_y_l<T, true>, (const void *)k_zinv_l<T>};
    for (const void *f : fns) {
        // dynamic LDS may use what the kernel's static LDS leaves of 160 KB
        hipFuncAttributes fa;
        hipError_t e = hipFuncGetAttributes(&fa, f);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, big - (int)fa.sharedSizeBytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" int voce_kernels_prepare(void)
{
    hipError_t e = set_lds_limits<double>();
    if (e == hipSuccess) e = set_lds_limits<float>();
    return (int)e;
}

// 2^a 3^b 5^c 7^d, 2..512 (planes beyond one block's LDS take the split
// plane passes; the LDS-tiled x pass holds a 512-point column)
extern "C" int voce_grid_supported(int n)
{
    if (n < 2 || n > SP_MAXN) return 0;
    int m = n;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    while (m % 7 == 0) m /= 7;
    return m == 1;
}

extern "C" int voce_vox_block(void) { return NT_VOX; }

// Default plane variant of an RVE before tuning: the line-parallel kernels
// where the grid allows them, else the 256- or 512-thread whole-plane
// kernels, else the split passes (line-parallel for 256 x 256 planes); -1 if
// none fits.
extern "C" int voce_default_plv(const RveDev *d)
{
    if (voce_plane_lds(d, 4) > 0) return 4;
    if (voce_plane_lds(d, 1) > 0) return 1;
    if (voce_plane_lds(d, 2) > 0) return 2;
    if (voce_plane_lds(d, 6) > 0) return 6;
    if (voce_plane_lds(d, 5) > 0) return 5;
    return -1;
}

static const int PL_MAXEL[4] = {MAXEL_PL_T, MAXEL_PL_S, MAXEL_PL_L, MAXEL_PL_M};
static const int X_MAXEL[3] = {MAXEL_X, MAXEL_XB, MAXEL_XC};
static const size_t LDS_MAX = 160 * 1024;

// Largest x-pass tile (columns per block, <= 16) for variant xv (variant 3:
// its smaller tile, two waves).
extern "C" int voce_xgreen_tile(int nx, int xv)
{
    if (xv == 3) return (nx == 16 || nx == 32) ? 4 : (nx == 64 ? 2 : 0);
    if (xv == 4) return (nx == 16 || nx == 32 || nx == 48 || nx == 64 || nx == 128) ? 8 : (nx == 256 ? 4 : 0);   // FP64: a whole column block (a half at 256)
    const int fit = X_MAXEL[xv] / (6 * nx);
    if (fit < 1) return 0;   // the variant cannot hold one column
    int tk = 1;
    while (2 * tk <= fit && 2 * tk <= 16) tk *= 2;
    return tk;               // the largest power of two <= 16 that fits
}

// Dynamic LDS of the plane and x-pass kernels of one RVE under its tile
// choices; 0 when a choice does not fit (tile larger than the variant's
// MAXEL or LDS beyond 160 KB).
// static LDS of the forward plane kernel's finaliser block, bounded above
static const size_t PLANE_STATIC_LDS = sizeof(MacroState) + 16 * 8 + (NT_PL_L / 64) * 14 * 8 + 256;

extern "C" size_t voce_plane_lds(const RveDev *r, int plv)
{
    const int el = r->ny * r->nzh;
    const size_t cb = r->fft_f32 ? 8 : 16;
    if (plv == 4) {
        // line-parallel variant: square planes of 16, 32, 48, 64 or 128, the plane only
        const int n = r->ny;
        if (r->nz != n || (n != 16 && n != 32 && n != 48 && n != 64 && n != 128)) return 0;
        const size_t b = (size_t)el * cb;
        return b + PLANE_STATIC_LDS <= LDS_MAX ? b : 0;
    }
    if (plv == 6) {
        // line-parallel split passes: 256 x 256 planes (their launches
        // take the kernels' own LDS sizes; this is the largest)
        if (r->ny != SL_N || r->nz != SL_N) return 0;
        const size_t b = r->fft_f32 ? std::max(sl_lds_z<float>(), sl_lds_y<float>()) : std::max(sl_lds_z<double>(), sl_lds_y<double>());
        return b + PLANE_STATIC_LDS <= LDS_MAX ? b : 0;
    }
    if (plv == 5) {
        // split z and y passes: a chunk of lines and one twiddle table
        if (r->ny > SP_MAXN || r->nz > SP_MAXN) return 0;
        const size_t b = (size_t)(MAXEL_SP + (r->ny > r->nz ? r->ny : r->nz)) * cb;
        return b + PLANE_STATIC_LDS <= LDS_MAX ? b : 0;
    }
    // an odd Ny packs its last row with a row of zeros: one more row
    const int elp = (r->ny + (r->ny & 1)) * r->nzh;
    if (plv < 0 || plv > 3 || elp > PL_MAXEL[plv]) return 0;
    const size_t b = (size_t)(elp + r->ny + r->nz) * cb;
    return b + PLANE_STATIC_LDS <= LDS_MAX ? b : 0;
}

extern "C" size_t voce_xgreen_lds(const RveDev *r, int xv, int tk)
{
    const size_t cb = r->fft_f32 ? 8 : 16;
    if (xv == 3) {
        // line-parallel x pass: Nx 16, 32 or 64; tk = 2 or 4 waves of CPW
        // columns; six lines of Nx per column
        const int nx = r->nx, cpw = nx == 64 ? 1 : 2;
        if (nx != 16 && nx != 32 && nx != 64) return 0;
        if (tk != 2 * cpw && tk != 4 * cpw) return 0;
        return (size_t)tk * 6 * nx * cb;
    }
    if (xv == 4) {
        // x pass staged through LDS: Nx 16, 32, 48, 64, 128 or 256; tk = a
        // whole column block (8 FP64 / 16 FP32 columns), a half or a quarter
        // of one (at 256 a half or a quarter); six lines of Nx per column
        const int nx = r->nx, tb = r->fft_f32 ? 16 : 8, cpw = nx >= 64 ? 1 : 2;
        if (nx != 16 && nx != 32 && nx != 48 && nx != 64 && nx != 128 && nx != 256) return 0;
        if ((tk != tb && 2 * tk != tb && 4 * tk != tb) || tk % cpw || (nx == 256 && tk == tb)) return 0;
        const size_t b = (size_t)tk * 6 * nx * cb;
        return b <= LDS_MAX ? b : 0;
    }
    if (xv < 0 || xv > 2 || tk < 1 || (tk & (tk - 1)) || 6 * tk * r->nx > X_MAXEL[xv]) return 0;
    // 6 spectra + 3 skew sums, lines padded to nx + 1, then the twiddles
    const size_t b = (size_t)(9 * tk * (r->nx + 1) + r->nx) * cb;
    return b <= LDS_MAX ? b : 0;
}

// Column-blocked spectrum layout of an RVE: ncb blocks of TB = 128 bytes of
// columns; nspec = complex elements per component (padded).
extern "C" void voce_spec_layout(RveDev *r)
{
    const int tb = r->fft_f32 ? 16 : 8;
    r->ncb = (r->ny * r->nzh + tb - 1) / tb;
    r->nspec = (long)r->ncb * tb * r->nx;
}

extern "C" int voce_launch_begin_increment(const LaunchSet *L, void *stream)
{
    hipLaunchKernelGGL(k_begin_increment, dim3(L->n_one), dim3(1), 0, (hipStream_t)stream, L->tab, L->map_one);
    return (int)hipGetLastError();
}

// One FFT kernel of the iteration alone (0: plane_fwd, 1: xgreen, 2:
// plane_inv), for timing launch choices.
extern "C" int voce_launch_fft_probe(const LaunchSet *L, void *stream, int which)
{
    hipStream_t st = (hipStream_t)stream;
    if (L->fft_f32) {
        if (which == 0) launch_plane_fwd<float>(L, st);
        else if (which == 1) launch_xgreen<float>(L, st, 0);
        else launch_plane_inv<float>(L, st, 0);
    } else {
        if (which == 0) launch_plane_fwd<double>(L, st);
        else if (which == 1) launch_xgreen<double>(L, st, 0);
        else launch_plane_inv<double>(L, st, 0);
    }
    return (int)hipGetLastError();
}

extern "C" int voce_launch_finalize(const LaunchSet *L, void *stream)
{
    hipLaunchKernelGGL((k_fin_iter<NT_FIN>), dim3(L->n_one), dim3(NT_FIN), 0, (hipStream_t)stream, L->tab, L->map_one);
    return (int)hipGetLastError();
}

extern "C" int voce_launch_iteration(const LaunchSet *L, void *stream, void *events)
{
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t *ev = (hipEvent_t *)events;
    return (int)(L->fft_f32 ? launch_iter_t<float>(L, st, ev) : launch_iter_t<double>(L, st, ev));
}

extern "C" int voce_launch_closeout(const LaunchSet *L, void *stream, int need_skew)
{
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (need_skew) e = L->fft_f32 ? launch_skew_t<float>(L, st) : launch_skew_t<double>(L, st);
    if (e != hipSuccess) return (int)e;
    if (L->uniform_phase) hipLaunchKernelGGL((k_closeout<NT_VOX, true>), dim3(L->n_vox), dim3(NT_VOX), 0, st, L->tab, L->map_vox);
    else hipLaunchKernelGGL((k_closeout<NT_VOX, false>), dim3(L->n_vox), dim3(NT_VOX), 0, st, L->tab, L->map_vox);
    hipLaunchKernelGGL((k_fin_incr<NT_FIN>), dim3(L->n_one), dim3(NT_FIN), 0, st, L->tab, L->map_one);
    return (int)hipGetLastError();
}

// Server poll: flags[2i] = active, flags[2i+1] = error of member idx[i].
__global__ __launch_bounds__(256) void k_poll(const RveDev *__restrict__ tab, const int *__restrict__ idx, int n, int *flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MacroState *ms = tab[idx[i]].ms;
    flags[2 * i] = ms->active;
    flags[2 * i + 1] = ms->error;
}

extern "C" int voce_launch_poll(const RveDev *tab, const int *idx_dev, int n, int *flags_dev, void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_poll, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, tab, idx_dev, n, flags_dev);
    return (int)hipGetLastError();
}

extern "C" int voce_launch_count_active(const RveDev *tab, int n, int *out_dev, void *stream)
{
    hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, (hipStream_t)stream, tab, n, out_dev);
    return (int)hipGetLastError();
}

extern "C" int voce_launch_init_fields(const RveDev *tab1, long V, const double *sig6_by_orient, void *stream)
{
    const long nb = (V + 255) / 256;
    hipLaunchKernelGGL(k_init_state, dim3(nb), dim3(256), 0, (hipStream_t)stream, tab1, sig6_by_orient);
    return (int)hipGetLastError();
}
