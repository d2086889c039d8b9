// fft_lds.hpp -- workgroup-cooperative FP64 complex FFTs on an LDS tile.
//
// NT butterflies per pass cover a tile of 4096 complex doubles in LDS
// (4096/NT points each: radix 16 for NT = 256, 8 for NT = 512), which holds
// 4096/FS independent transforms of size FS; a smaller radix takes the one
// mixed pass when log2(FS) is not a multiple of the main radix's log.  Data
// goes LDS -> registers -> butterfly -> LDS, two barriers per pass; the IO
// variant feeds the first pass from a loader and drains the last into a
// storer.  Layouts (below): SwzLayout for radix-8 passes, Swz16Layout for
// radix-16 passes, ColLayout for the four-step column kernels.
//
// Twiddles come from host tables computed from long-double cos/sin (one
// rounding per entry), no on-device sin/cos.  Table layout (see tw_big):
//   tw[m],        m < 4096:  exp(-2 pi i m / 4096)     (all in-tile passes)
//   tw[4096 + f], f < FT/4096: exp(-2 pi i f / FT)     (fine part, FT = 2^twShift)
// In-tile passes index the dense 4096-entry part, so their loads share cache
// lines; four-step twiddles w_F^m are one coarse entry times one fine entry.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace icm {

constexpr int kTile = 4096;   // complex points per workgroup tile

// w_F^m for F = 2^logF <= FT = 2^twShift and 0 <= m < F: M = m FT / F split
// into a coarse index (top 12 bits) and a fine index; one complex product.
__device__ __forceinline__ double2 tw_big(const double2 *tw, int twShift, int m, int logF) {
    const int M = m << (twShift - logF);
    const int lf = twShift - 12;
    const double2 c = tw[M >> lf];
    if (lf <= 0) return c;
    const double2 f = tw[kTile + (M & ((1 << lf) - 1))];
    return make_double2(c.x * f.x - c.y * f.y, c.x * f.y + c.y * f.x);
}

// tw_big for transforms larger than the dense table (logF > 12, so the
// table always has its fine part): no branch, so the two loads of every
// output issue together.  It costs registers where many outputs are in
// flight, so callers choose it by measurement.
__device__ __forceinline__ double2 tw_big2(const double2 *tw, int twShift, int m, int logF) {
    const int M = m << (twShift - logF);
    const int lf = twShift - 12;
    const double2 c = tw[M >> lf];
    const double2 f = tw[kTile + (M & ((1 << lf) - 1))];
    return make_double2(c.x * f.x - c.y * f.y, c.x * f.y + c.y * f.x);
}

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

// Unpadded, XOR-swizzled tile for the 512-thread (radix-8) kernels, whose
// accesses are always runs of consecutive elements of one transform or
// Stockham writes with stride R*Ns.  Measured on gfx950 (LDS probe,
// SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE): ds_write_b128 is conflict-free
// when each group of 8 lanes hits 8 distinct 16-byte units mod 8, and
// ds_read_b128 when each group of 16 lanes hits 16 distinct units mod 16;
// padding a contiguous run (idx + idx/16) makes the reads conflict 2-way.
// Element idx of transform it lives at it*FS + (idx ^ ((idx >> 3) & 7) ^ g(it)):
//   * the low-3-bit XOR by idx bits 3..5 turns the stride-2/4/8 writes of the
//     first passes into 8 distinct units and permutes runs only inside their
//     aligned 8-blocks (contiguous reads stay conflict-free);
//   * g(it) spreads the transforms of FS = 16, 32, 64, where one 8- or
//     16-lane group spans 2..8 transforms (the passes with FS/R < 16).
template <int FS> struct SwzLayout {
    static constexpr bool kItemFast = false;
    static constexpr int IS = FS;
    static constexpr int ITEMS = kTile / FS;
    static constexpr int SIZE = kTile;
    __host__ __device__ static __forceinline__ int g(int item) {
        if constexpr (FS == 16) {
            const int b0 = item & 1, b1 = (item >> 1) & 1, b2 = (item >> 2) & 1;
            return ((b0 ^ b2) << 3) | (b1 << 2) | (b0 << 1);
        } else if constexpr (FS == 32) {
            const int b0 = item & 1, b1 = (item >> 1) & 1;
            return ((b0 ^ b1) << 3) | (b0 << 2);
        } else if constexpr (FS == 64) {
            return (item & 1) << 3;
        } else {
            return 0;
        }
    }
    __host__ __device__ static __forceinline__ int at(int item, int idx) {
        return item * FS + (idx ^ ((idx >> 3) & 7) ^ g(item));
    }
};

// Unpadded, XOR-swizzled tile for radix-16 passes (256 butterflies per pass)
// over transforms of FS = 128 .. 4096 points.  Under the gfx950 LDS rules
// above: a first radix-16 pass writes with stride 16 (idx = 16 l + o over
// lanes l), which an XOR of index bits 4..6 into the low three bits spreads
// over 8 distinct 16-byte units per 8-lane group; when log2(FS) is not a
// multiple of 4, the first pass is a radix-2/4/8 one writing with stride
// 2/4/8, and the XOR takes bits 3..5 as well.  Later passes write, and every
// pass reads, runs of consecutive elements, which such XORs only permute
// inside their aligned 8-blocks.  g(it) flips bit 3 for odd transforms of
// FS = 128, where one 16-lane read group spans two transforms.
template <int FS> struct Swz16Layout {
    static_assert(FS >= 128 && FS <= kTile, "radix-16 tiles hold transforms of 128 .. 4096 points");
    static constexpr bool kItemFast = false;
    static constexpr int IS = FS;
    static constexpr int ITEMS = kTile / FS;
    static constexpr int SIZE = kTile;
    static constexpr bool kMixed = (ilog2c(FS) % 4) != 0;
    __host__ __device__ static __forceinline__ int at(int item, int idx) {
        if constexpr (kMixed) {
            const int g = (FS == 128) ? ((item & 1) << 3) : 0;
            return item * FS + (idx ^ (((idx >> 3) ^ (idx >> 4)) & 7) ^ g);
        } else {
            return item * FS + (idx ^ ((idx >> 4) & 7));
        }
    }
};

// Row-major tile for the column kernels: element idx of transform it at
// idx * ITEMS + it, and the passes give consecutive lanes consecutive
// transforms at the same butterfly (kItemFast).  Loads, stores and every
// pass read / write then touch runs of consecutive elements: no bank
// conflicts and no padding while ITEMS >= 8 (FS <= 512); wider columns
// (FS = 1024 .. 4096, 4 .. 1 per tile) stay correct with 2..8-way write
// conflicts in their first passes.
template <int FS> struct ColLayout {
    static constexpr bool kItemFast = true;
    static constexpr int ITEMS = kTile / FS;
    static constexpr int IS = 1;
    static constexpr int SIZE = kTile;
    __host__ __device__ static __forceinline__ int at(int item, int idx) { return idx * ITEMS + item; }
};

// The thread index through an empty asm statement: every pass then derives
// its LDS addresses afresh instead of the compiler computing all passes'
// addresses once and keeping them live (in registers) across the transform,
// which pushes the 128-VGPR kernels into spills.  No instruction is emitted.
// KEEP = true leaves the choice to the compiler (kernels that run one
// workgroup per CU and have registers to spare measure faster that way).
template <bool KEEP>
__device__ __forceinline__ int pass_tid() {
    int t = threadIdx.x;
    if constexpr (!KEEP) asm volatile("" : "+v"(t));
    return t;
}

__device__ __forceinline__ double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 c_mulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 c_conj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 c_scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// Multiply by the forward (-i) or inverse (+i) quarter turn.
template <bool INV> __device__ __forceinline__ double2 mul_q(double2 a) {
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
// Multiply by the constant (cr, ci) for the forward direction; conjugate for inverse.
template <bool INV> __device__ __forceinline__ double2 mul_k(double2 a, double cr, double ci) {
    const double c_i = INV ? -ci : ci;
    return make_double2(a.x * cr - a.y * c_i, a.x * c_i + a.y * cr);
}

// cos(pi/8), sin(pi/8), sqrt(2)/2 rounded to FP64.
constexpr double kC8 = 0.92387953251128675613;
constexpr double kS8 = 0.38268343236508977173;
constexpr double kR2 = 0.70710678118654752440;

template <bool INV> __device__ __forceinline__ void dft2(double2 &a0, double2 &a1) {
    double2 t = a0;
    a0 = c_add(t, a1);
    a1 = c_sub(t, a1);
}

template <bool INV> __device__ __forceinline__ void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3) {
    double2 t0 = c_add(a0, a2), t1 = c_sub(a0, a2);
    double2 t2 = c_add(a1, a3), t3 = mul_q<INV>(c_sub(a1, a3));
    a0 = c_add(t0, t2);
    a2 = c_sub(t0, t2);
    a1 = c_add(t1, t3);
    a3 = c_sub(t1, t3);
}

template <bool INV> __device__ __forceinline__ void dft8(double2 *a) {
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    a[3] = mul_k<INV>(a[3], kR2, -kR2);
    a[5] = mul_q<INV>(a[5]);
    a[7] = mul_k<INV>(a[7], -kR2, -kR2);
    double2 x0 = c_add(a[0], a[1]), x4 = c_sub(a[0], a[1]);
    double2 x1 = c_add(a[2], a[3]), x5 = c_sub(a[2], a[3]);
    double2 x2 = c_add(a[4], a[5]), x6 = c_sub(a[4], a[5]);
    double2 x3 = c_add(a[6], a[7]), x7 = c_sub(a[6], a[7]);
    a[0] = x0; a[1] = x1; a[2] = x2; a[3] = x3;
    a[4] = x4; a[5] = x5; a[6] = x6; a[7] = x7;
}

// 16-point DFT as 4 x 4: X[k1 + 4 k2] = sum_s w4^(s k2) w16^(s k1) DFT4_u(a[s + 4u])[k1].
// Output is left in a[4 k1 + k2] (digit-reversed); callers map register j to
// output index out16(j) = j / 4 + 4 (j % 4) when storing.
template <bool INV> __device__ __forceinline__ void dft16(double2 *a) {
#pragma unroll
    for (int s = 0; s < 4; ++s) dft4<INV>(a[s], a[s + 4], a[s + 8], a[s + 12]);
    // a[s + 4 k1] *= w16^(s k1)
    a[5]  = mul_k<INV>(a[5],  kC8, -kS8);    // s=1,k1=1: w1
    a[9]  = mul_k<INV>(a[9],  kR2, -kR2);    // s=1,k1=2: w2
    a[13] = mul_k<INV>(a[13], kS8, -kC8);    // s=1,k1=3: w3
    a[6]  = mul_k<INV>(a[6],  kR2, -kR2);    // s=2,k1=1: w2
    a[10] = mul_q<INV>(a[10]);               // s=2,k1=2: w4
    a[14] = mul_k<INV>(a[14], -kR2, -kR2);   // s=2,k1=3: w6
    a[7]  = mul_k<INV>(a[7],  kS8, -kC8);    // s=3,k1=1: w3
    a[11] = mul_k<INV>(a[11], -kR2, -kR2);   // s=3,k1=2: w6
    a[15] = mul_k<INV>(a[15], -kC8, kS8);    // s=3,k1=3: w9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<INV>(a[4 * k1], a[4 * k1 + 1], a[4 * k1 + 2], a[4 * k1 + 3]);
}

// Output index held by register j after dft<R>.
template <int R> __device__ __forceinline__ constexpr int dft_out(int j) {
    return R == 16 ? (j >> 2) + 4 * (j & 3) : j;
}

template <int R, bool INV> __device__ __forceinline__ void dft(double2 *a) {
    if constexpr (R == 2) dft2<INV>(a[0], a[1]);
    else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 8) dft8<INV>(a);
    else dft16<INV>(a);
}


// v[s] *= w^s (s = 1..R-1), w = exp(-+2 pi i k / (Ns R)).  Only w, w^2, w^4
// and w^8 come from the table (correctly rounded); the other powers are
// products of at most three of them, so each twiddle carries a few ulp at
// most.  Loading 4 instead of 15 entries keeps the register footprint small.
// The passes load the entries (twiddle_load) before their LDS reads and
// barrier and apply them after (twiddle_apply), so the table's latency
// overlaps the reads and the wait at the barrier.  The table pointer is
// deliberately not __restrict__: its loads then stay below the preceding
// pass's barrier instead of being hoisted across passes.
template <int R> struct TwRaw {
    double2 e[R >= 16 ? 4 : R >= 8 ? 3 : R >= 4 ? 2 : 1];   // w, w^2, w^4, w^8
};
template <int R>
__device__ __forceinline__ TwRaw<R> twiddle_load(const double2 *tw, int k, int twSh) {
    TwRaw<R> r;
    r.e[0] = tw[k << twSh];
    if constexpr (R >= 4) r.e[1] = tw[(2 * k) << twSh];
    if constexpr (R >= 8) r.e[2] = tw[(4 * k) << twSh];
    if constexpr (R >= 16) r.e[3] = tw[(8 * k) << twSh];
    return r;
}
template <int R, bool INV>
__device__ __forceinline__ void twiddle_apply(double2 *v, const TwRaw<R> &r) {
    double2 w[R];
    w[1] = r.e[0];
    if constexpr (R >= 4) w[2] = r.e[1];
    if constexpr (R >= 8) w[4] = r.e[2];
    if constexpr (R >= 16) w[8] = r.e[3];
    if constexpr (R >= 4) w[3] = c_mul(w[1], w[2]);
    if constexpr (R >= 8) {
        w[5] = c_mul(w[1], w[4]);
        w[6] = c_mul(w[2], w[4]);
        w[7] = c_mul(w[3], w[4]);
    }
    if constexpr (R >= 16) {
#pragma unroll
        for (int s = 9; s < 16; ++s) w[s] = c_mul(w[s - 8], w[8]);
    }
#pragma unroll
    for (int s = 1; s < R; ++s) v[s] = INV ? c_mulc(v[s], w[s]) : c_mul(v[s], w[s]);
}

// One Stockham pass of radix R over all transforms of size FS in the tile,
// executed by NT threads (each owning kTile / NT points).
// Ns = product of the radices of earlier passes.
template <int FS, int R, bool INV, int NT, class TL>
__device__ __forceinline__ void fft_pass(double2 *lds, int logNs, const double2 *tw, int twShift) {
    constexpr int NBI = FS / R;               // butterflies per transform
    constexpr int LOG_NBI = ilog2c(NBI);
    constexpr int PER = (kTile / NT) / R;     // butterflies per thread
    const int t = pass_tid<false>();
    const int Ns = 1 << logNs;
    double2 v[PER][R];
    int item[PER], jj[PER];
    TwRaw<R> w[PER];
    const int twSh = 12 - (logNs + ilog2c(R));   // into the 4096-entry part
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int bj = t + u * NT;
        if constexpr (TL::kItemFast) {
            constexpr int LOG_ITEMS = ilog2c(kTile / FS);
            item[u] = bj & ((kTile / FS) - 1);
            jj[u] = bj >> LOG_ITEMS;
        } else {
            item[u] = bj >> LOG_NBI;
            jj[u] = bj & (NBI - 1);
        }
        if (logNs > 0) w[u] = twiddle_load<R>(tw, jj[u] & (Ns - 1), twSh);
#pragma unroll
        for (int s = 0; s < R; ++s) v[u][s] = lds[TL::at(item[u], jj[u] + s * NBI)];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int k = jj[u] & (Ns - 1);
        if (logNs > 0) twiddle_apply<R, INV>(v[u], w[u]);
        dft<R, INV>(v[u]);
        const int d0 = (jj[u] - k) * R + k;
#pragma unroll
        for (int j = 0; j < R; ++j) lds[TL::at(item[u], d0 + dft_out<R>(j) * Ns)] = v[u][j];
    }
    __syncthreads();
}

// Transforms whose first pass reads its inputs through a loader (global
// memory -> registers, no LDS store + load before the first butterflies)
// and/or whose last pass hands its outputs to a storer (no LDS store + load
// after the last butterflies).  Ld: double2 operator()(int item, int idx);
// St: void operator()(int item, int idx, double2 v); NoIO disables either.
struct NoIO {};

template <int FS, int R, bool INV, int NT, class TL, int PTS = kTile, bool KEEP = false, class Ld, class St>
__device__ __forceinline__ void fft_pass_io(double2 *lds, int logNs, const double2 *tw, const Ld &ld, const St &st) {
    constexpr bool LOAD = !std::is_same<Ld, NoIO>::value;
    constexpr bool STORE = !std::is_same<St, NoIO>::value;
    constexpr int NBI = FS / R;
    constexpr int LOG_NBI = ilog2c(NBI);
    constexpr int PER = (PTS / NT) / R;
    static_assert(PTS == kTile || !TL::kItemFast, "item-fast layouts hold one tile");
    const int t = pass_tid<KEEP>();
    const int Ns = 1 << logNs;
    double2 v[PER][R];
    int item[PER], jj[PER];
    TwRaw<R> w[PER];
    const int twSh = 12 - (logNs + ilog2c(R));
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int bj = t + u * NT;
        if constexpr (TL::kItemFast) {
            constexpr int LOG_ITEMS = ilog2c(kTile / FS);
            item[u] = bj & ((kTile / FS) - 1);
            jj[u] = bj >> LOG_ITEMS;
        } else {
            item[u] = bj >> LOG_NBI;
            jj[u] = bj & (NBI - 1);
        }
        if (logNs > 0) w[u] = twiddle_load<R>(tw, jj[u] & (Ns - 1), twSh);
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if constexpr (LOAD) v[u][s] = ld(item[u], jj[u] + s * NBI);
            else v[u][s] = lds[TL::at(item[u], jj[u] + s * NBI)];
        }
    }
    // loads: the caller's earlier LDS reads must finish before this pass
    // writes; LDS reads: every read of this pass before anyone writes
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int k = jj[u] & (Ns - 1);
        if (logNs > 0) twiddle_apply<R, INV>(v[u], w[u]);
        dft<R, INV>(v[u]);
        const int d0 = (jj[u] - k) * R + k;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if constexpr (STORE) st(item[u], d0 + dft_out<R>(j) * Ns, v[u][j]);
            else lds[TL::at(item[u], d0 + dft_out<R>(j) * Ns)] = v[u][j];
        }
    }
    if constexpr (!STORE) __syncthreads();
}

// fft_tile with the first pass fed by `ld` and / or the last pass emptied
// into `st` (NoIO: through LDS as fft_tile).  With a storer the tile is free
// for writing when the call returns (the last pass's read barrier).
// PTS = points transformed per call: kTile (one tile), or a multiple of it,
// in which case every thread runs PTS / kTile butterflies per pass at the
// same index of consecutive tiles (the radix stays kTile / NT).
template <int FS, bool INV, int NT, class TL, int PTS = kTile, bool KEEP = false, class Ld, class St>
__device__ __forceinline__ void fft_tile_io(double2 *lds, const double2 *tw, const Ld &ld, const St &st) {
    constexpr int RMAX = kTile / NT;
    constexpr int LR = ilog2c(RMAX);
    constexpr int LOG = ilog2c(FS);
    constexpr int REM = LOG % LR;
    constexpr int NP = (REM ? 1 : 0) + LOG / LR;   // passes
    static_assert(NP >= 2, "first and last pass must differ");
    if constexpr (std::is_same<Ld, NoIO>::value) __syncthreads();
    int logNs = 0;
    if constexpr (REM == 1) { fft_pass_io<FS, 2, INV, NT, TL, PTS, KEEP>(lds, logNs, tw, ld, NoIO{}); logNs += 1; }
    if constexpr (REM == 2) { fft_pass_io<FS, 4, INV, NT, TL, PTS, KEEP>(lds, logNs, tw, ld, NoIO{}); logNs += 2; }
    if constexpr (REM == 3) { fft_pass_io<FS, 8, INV, NT, TL, PTS, KEEP>(lds, logNs, tw, ld, NoIO{}); logNs += 3; }
#pragma unroll
    for (int p = 0; p < LOG / LR; ++p) {
        const bool first = (REM == 0 && p == 0), last = (p == LOG / LR - 1);
        if (first && last) {
            fft_pass_io<FS, RMAX, INV, NT, TL, PTS, KEEP>(lds, logNs, tw, ld, st);
        } else if (first) {
            fft_pass_io<FS, RMAX, INV, NT, TL, PTS, KEEP>(lds, logNs, tw, ld, NoIO{});
        } else if (last) {
            fft_pass_io<FS, RMAX, INV, NT, TL, PTS, KEEP>(lds, logNs, tw, NoIO{}, st);
        } else {
            fft_pass_io<FS, RMAX, INV, NT, TL, PTS, KEEP>(lds, logNs, tw, NoIO{}, NoIO{});
        }
        logNs += LR;
    }
}

// In-place transform of every size-FS transform of the tile (layout TL) by NT
// threads (256: radix-16 passes; 512: radix-8 passes), after one mixed pass
// when log2(FS) is not a multiple of the main radix's log.
// Starts with a barrier so callers may write the tile and call directly;
// ends with a barrier (the last pass's).
template <int FS, bool INV, int NT, class TL>
__device__ __forceinline__ void fft_tile(double2 *lds, const double2 *tw, int twShift) {
    constexpr int RMAX = kTile / NT;          // 16 or 8
    constexpr int LR = ilog2c(RMAX);
    constexpr int LOG = ilog2c(FS);
    constexpr int REM = LOG % LR;
    __syncthreads();
    int logNs = 0;
    if constexpr (REM == 1) { fft_pass<FS, 2, INV, NT, TL>(lds, logNs, tw, twShift); logNs += 1; }
    if constexpr (REM == 2) { fft_pass<FS, 4, INV, NT, TL>(lds, logNs, tw, twShift); logNs += 2; }
    if constexpr (REM == 3) { fft_pass<FS, 8, INV, NT, TL>(lds, logNs, tw, twShift); logNs += 3; }
#pragma unroll
    for (int p = 0; p < LOG / LR; ++p) {
        fft_pass<FS, RMAX, INV, NT, TL>(lds, logNs, tw, twShift);
        logNs += LR;
    }
}

}  // namespace icm
