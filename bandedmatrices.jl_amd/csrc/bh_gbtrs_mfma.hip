// bh_gbtrs_mfma.hip -- blocked band solve with the factors of bh_?gbtrf,
// panels of PB rows applied on the matrix cores (gfx950 MFMA 16x16x4).
//
// The reference solves with LAPACK.gbtrs! (src/banded/linalg.jl:24-30):
// xGBTRS applies, column by column, the row interchange of step j and the
// rank-1 update of the rows below it (the forward sweep with L), then the
// upper band solve with U (xTBSV).  Both sweeps are a chain of n dependent
// steps.  Here the chain has n/PB links instead:
//
// Forward.  For panel P (rows j0 = P*PB .. j0+PB-1) the window is the RW =
// kl+PB rows its interchanges and multipliers touch.  Applying the PB steps
// of the panel to the window equals (with the panel's row permutation perm,
// its multipliers in final row order L11 (unit lower, PB x PB) and L21
// (kl x PB))
//     Y1 = L11^{-1} (perm W)[0:PB]            the panel's finished rows
//     W' = (perm W)[PB:RW] - L21 Y1           the rows the next panel sees
//        = (perm W)[PB:RW] - M21 (perm W)[0:PB],   M21 = L21 L11^{-1}
// perm, L11^{-1} and M21 depend on the factors only; a parallel pre-pass
// (one workgroup per panel) builds them once per factorisation.  L21 in
// final order is xGBTRF's form of the multipliers before it "partially
// undoes" the interchanges: each stored column c is moved to the positions
// its rows reach after the panel's later interchanges.
// Backward.  Panels from the last: X1 = U11^{-1} Y[j0:j0+PB], and the kv =
// kl+ku rows above the panel lose U[j0-kv:j0, j0:j0+PB] X1 =
// V Y[j0:j0+PB], V = U[j0-kv:j0, j0:j0+PB] U11^{-1}.
//
// Both products of a link read only the window as it stands, so they run
// side by side: one workgroup solves 16 right-hand sides with one MFMA
// product per 16-row tile and one LDS-only barrier per link; the operands
// are loaded two links ahead.  Results equal xGBTRS's to rounding (the
// diagonal blocks through their inverses, a panel's updates in one sum),
// not bit for bit: this solve serves the mixed-precision refinement, whose
// corrections need an accurate solve, not LAPACK's operation order;
// bh_?gbtrs stays the bit-exact path.
#include <type_traits>

#include "bh_common.hpp"
#include "bh_mfma.hpp"

namespace bh {

namespace {

template <typename T>
struct TStride {
    static constexpr int value = sizeof(T) == 4 ? 20 : 18;  // LDS row stride of [rows][16] tiles
};

// Per panel the pre-pass writes one record of A operands, each 16-row tile
// lane-major (lane l's PB/4 K-chunk values contiguous, so a lane loads its
// operands with 16-byte reads): L11^{-1} (PB/16 tiles), M21 (KT tiles),
// U11^{-1} (PB/16 tiles), -V (KTU tiles) -- the update blocks negated, so
// the sweeps accumulate onto the rows' current values; and one record of
// 16-bit window rows, as byte offsets into the forward sweep's LDS window:
// the gather rows of the diagonal-block product (4 x PB/4) and the C-operand
// rows of each M21 tile (KT x 16).
template <typename T, int PB>
struct TrsLayout {
    int n, kl, ku, kv, RW, KT, KTU, np;
    size_t linv_off, l21_off, uinv_off, ublk_off, per_panel;  // elements of T
    size_t cidx_off, idx_u16;                                 // 16-bit entries
    static TrsLayout make(i64 n, i64 kl, i64 ku) {
        TrsLayout L{};
        L.n = (int)n;
        L.kl = (int)kl;
        L.ku = (int)ku;
        L.kv = (int)(kl + ku);
        L.RW = (int)kl + PB;
        L.KT = (int)((kl + 15) / 16);
        L.KTU = (int)((kl + ku + 15) / 16);
        L.np = (int)((n + PB - 1) / PB);
        L.linv_off = 0;
        L.l21_off = (size_t)PB * PB;
        L.uinv_off = L.l21_off + (size_t)L.KT * 16 * PB;
        L.ublk_off = L.uinv_off + (size_t)PB * PB;
        L.per_panel = L.ublk_off + (size_t)L.KTU * 16 * PB;
        L.cidx_off = (size_t)PB;  // 4 x PB/4 gather rows first
        L.idx_u16 = (L.cidx_off + (size_t)L.KT * 16 + 7) & ~(size_t)7;
        return L;
    }
};

// ---------------------------------------------------------------- pre-pass
// One workgroup per panel: perm, L11^{-1} and L21 in final row order, U11^{-1},
// the matrices stored in MFMA A-operand order (16x4 blocks, lane i+16k holding
// element (i, k)), so the solve kernels load them with one coalesced read each.
template <typename T, int PB>
__global__ __launch_bounds__(256) void trs_prep(const T *__restrict__ ab, i64 ldab, const i64 *__restrict__ ipiv,
                                                TrsLayout<T, PB> L, unsigned short *__restrict__ idxA,
                                                T *__restrict__ rec) {
    extern __shared__ __align__(16) unsigned char tp_raw[];
    constexpr int LS = PB + 1;
    const int RW = L.RW, n = L.n, kl = L.kl, kv = L.kv;
    const int top = 16 * L.KTU;
    T *Lf = reinterpret_cast<T *>(tp_raw);  // RW x LS
    T *Ut = Lf + (size_t)RW * LS;           // PB x LS
    T *Xi = Ut + PB * LS;                   // PB x LS
    T *Ub = Xi + PB * LS;                   // top x LS: U block above the panel
    int *perm = reinterpret_cast<int *>(Ub + (size_t)top * LS);
    int *F = perm + RW;
    __shared__ int q[PB];
    const int p = blockIdx.x, tid = threadIdx.x;
    const int j0 = p * PB, nb = min(PB, n - j0);
    for (int e = tid; e < RW * LS; e += 256) Lf[e] = T(0);
    for (int r = tid; r < RW; r += 256) {
        perm[r] = r;
        F[r] = r;
    }
    if (tid < PB) {
        int qq = tid;
        if (tid < nb) {
            const i64 piv = ipiv[j0 + tid] - 1 - j0;  // 0 <= piv - tid <= kl for valid factors
            qq = (piv >= tid && piv < RW) ? (int)piv : tid;
        }
        q[tid] = qq;
    }
    __syncthreads();
    if (tid == 0) {
        for (int c = 0; c < nb; ++c) {
            const int a = perm[c];
            perm[c] = perm[q[c]];
            perm[q[c]] = a;
        }
    }
    // multipliers of column c sit at positions c+1..c+kl after step c; the
    // later interchanges of the panel carry position r to F_c(r)
    for (int c = nb - 1; c >= 0; --c) {
        const i64 col = (i64)(j0 + c) * ldab;
        for (int r = c + 1 + tid; r <= c + kl && r < RW; r += 256)
            if (j0 + r < n) Lf[F[r] * LS + c] = ab[col + (kv + r - c)];
        __syncthreads();
        if (tid == 0 && q[c] != c) {
            const int a = F[c];
            F[c] = F[q[c]];
            F[q[c]] = a;
        }
        __syncthreads();
    }
    if (tid < PB) Lf[tid * LS + tid] = T(1);
    // U11 (upper, identity past the last column)
    for (int e = tid; e < PB * PB; e += 256) {
        const int i = e / PB, j = e - i * PB;
        T v = T(0);
        if (j >= nb)
            v = (i == j) ? T(1) : T(0);
        else if (i <= j && j - i <= kv)
            v = ab[(i64)(j0 + j) * ldab + (kv + i - j)];
        Ut[i * LS + j] = v;
    }
    __syncthreads();
    // L11^{-1}: column j by forward substitution (thread j)
    if (tid < PB) {
        const int j = tid;
        for (int i = 0; i < PB; ++i) {
            T x = (i == j) ? T(1) : T(0);
            if (i > j) {
                T s = T(0);
                for (int k = j; k < i; ++k) s = dfma(Lf[i * LS + k], Xi[k * LS + j], s);
                x = -s;
            }
            Xi[i * LS + j] = x;
        }
    }
    __syncthreads();
    constexpr int NS = PB / 4;
    T *out = rec + (size_t)p * L.per_panel;
    // element e of a tile block: tile e / (64 NS), lane (e / NS) % 64, chunk e % NS
    for (int e = tid; e < PB * PB; e += 256) {
        const int s = e % NS, lane = (e / NS) & 63, ti = e / (64 * NS);
        out[L.linv_off + e] = Xi[(16 * ti + (lane & 15)) * LS + 4 * s + (lane >> 4)];
    }
    // -M21 = -L21 L11^{-1} (L11^{-1} lower triangular)
    for (int e = tid; e < L.KT * 16 * PB; e += 256) {
        const int s = e % NS, lane = (e / NS) & 63, t = e / (64 * NS);
        const int r = PB + 16 * t + (lane & 15), c = 4 * s + (lane >> 4);
        T acc = T(0);
        if (r < RW)
            for (int k = c; k < PB; ++k) acc = dfma(Lf[r * LS + k], Xi[k * LS + c], acc);
        out[L.l21_off + e] = -acc;  // stored negated: the sweep accumulates onto the C operand
    }
    // U(i, j) above the panel: i = j0 - top + r, j = j0 + c
    for (int e = tid; e < top * PB; e += 256) {
        const int r = e / PB, c = e - r * PB;
        const int i = j0 - top + r, j = j0 + c, d = kv + i - j;
        Ub[r * LS + c] = (i >= 0 && j < n && d >= 0) ? ab[(i64)j * ldab + d] : T(0);
    }
    unsigned short *ix = idxA + (size_t)p * L.idx_u16;
    // window rows as LDS byte offsets of the sweep's [rows][SW] buffer (< 2^16)
    constexpr int RB_ = TStride<T>::value * (int)sizeof(T);
    for (int e = tid; e < PB; e += 256) ix[e] = (unsigned short)(perm[4 * (e % NS) + e / NS] * RB_);  // [kq][s]
    for (int e = tid; e < L.KT * 16; e += 256) {
        const int v = e & 3, kq = (e >> 2) & 3, t = e >> 4;
        ix[L.cidx_off + e] = (unsigned short)(perm[min(PB + 16 * t + drow<T>(kq, v), RW - 1)] * RB_);
    }
    __syncthreads();
    // U11^{-1}: column j by back substitution (thread j)
    if (tid < PB) {
        const int j = tid;
        for (int i = PB - 1; i >= 0; --i) {
            T x = T(0);
            if (i <= j) {
                T s = (i == j) ? T(1) : T(0);
                for (int k = i + 1; k <= j; ++k) s = dfma(-Ut[i * LS + k], Xi[k * LS + j], s);
                x = s / Ut[i * LS + i];
            }
            Xi[i * LS + j] = x;
        }
    }
    __syncthreads();
    for (int e = tid; e < PB * PB; e += 256) {
        const int s = e % NS, lane = (e / NS) & 63, ti = e / (64 * NS);
        out[L.uinv_off + e] = Xi[(16 * ti + (lane & 15)) * LS + 4 * s + (lane >> 4)];
    }
    // V = U_block U11^{-1} (U11^{-1} upper triangular), stored negated
    for (int e = tid; e < L.KTU * 16 * PB; e += 256) {
        const int s = e % NS, lane = (e / NS) & 63, t = e / (64 * NS);
        const int r = 16 * t + (lane & 15), c = 4 * s + (lane >> 4);
        T acc = T(0);
        for (int k = 0; k <= c; ++k) acc = dfma(Ub[r * LS + k], Xi[k * LS + c], acc);
        out[L.ublk_off + e] = -acc;
    }
}

// Operands of one panel, loaded two panels ahead of their use (three sets
// rotate), so the load latency is hidden behind two panels of work.  Every
// wave issues the same buffer loads and stores in every panel (clamped
// indices; out-of-range offsets read 0 and drop stores), so the compiler's
// wait counts stay exact: a panel waits only for loads two panels old.
template <typename T, int NS, int MT>
struct PanelOps {
    T a[NS];       // A operands of the diagonal-block product
    T b[MT][NS];   // A operands of this wave's update tiles
    int g[NS];     // forward: window rows gathered as B operands (rows 4s+kq)
    int c[MT][4];  // forward: window rows of the update tiles' C operands
    T fresh;       // one element of the rows entering the window
};

typedef unsigned trs_v4u __attribute__((ext_vector_type(4)));
typedef unsigned trs_v2u __attribute__((ext_vector_type(2)));

// N consecutive elements at byte offset off (16-byte aligned) with 16-byte loads
template <typename T, int N>
__device__ __forceinline__ void bload_vec(__amdgpu_buffer_rsrc_t r, int off, T (&dst)[N]) {
    static_assert((N * sizeof(T)) % 16 == 0, "whole 16-byte loads");
#pragma unroll
    for (int q = 0; q < (int)(N * sizeof(T) / 16); ++q) {
        const trs_v4u x = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * q, 0, 0);
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // through a scalar: __builtin_bit_cast of a vector element
                // (x[i] directly) is lowered wrongly by this compiler
                const unsigned e = x[i];
                dst[4 * q + i] = __builtin_bit_cast(float, e);
            }
        } else {
            dst[2 * q] = __builtin_bit_cast(double, (trs_v2u){x[0], x[1]});
            dst[2 * q + 1] = __builtin_bit_cast(double, (trs_v2u){x[2], x[3]});
        }
    }
}

// ---------------------------------------------------------------- forward sweep
// B (n x nrhs, ldb) <- L^{-1} P B for the 16 columns starting at 16*blockIdx.x.
// LDS addresses are precomputed per lane; the window buffer a panel reads
// (cur) is a compile-time constant of each of the six unrolled panels.
template <typename T, int PB, int NW, int MT>
__global__ __launch_bounds__(64 * NW) void trs_fwd(T *B, i64 ldb, int nrhs, TrsLayout<T, PB> L,
                                                   const unsigned short *idxA, const T *rec) {
    using v4 = typename Mf<T>::v4;
    constexpr int SW = TStride<T>::value;
    constexpr int NS = PB / 4;   // K chunks of a panel product
    constexpr int NP = PB / 16;  // 16-row tiles of a panel (MT: M21 tiles per wave)
    constexpr int ES = (int)sizeof(T);
    constexpr int RS = SW * ES;  // LDS bytes per window row
    static_assert(NS == 8, "the gather rows of a lane are one 16-byte load");
    using Ops = PanelOps<T, NS, MT>;
    extern __shared__ __align__(16) unsigned char tf_raw[];
    const int RW = L.RW, KT = L.KT, n = L.n, np = L.np;
    const int RWP = 16 * ((RW + 15) / 16) + 1;  // + one dump row for results past RW
    const int WB = RWP * RS;                    // bytes of one window buffer
    // w is wave-uniform (branches on it stay scalar)
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int jc = lane & 15, kq = lane >> 4;
    const int fr = tid >> 4, fj = tid & 15;  // PB*16 <= 64*NW threads carry one fresh element
    const int c0 = 16 * blockIdx.x;
    const bool colok = c0 + jc < nrhs;
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(B + (i64)c0 * ldb, 0, (int)(16 * ldb * ES),
                                                                        kQRsrcWord3);
    // per-lane byte offsets inside a panel's records
    const int offA = (int)((L.linv_off + ((size_t)min(w, NP - 1) * 64 + lane) * NS) * ES);
    int offB[MT], offC[MT], wdst[MT][4];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int t = min(w + m * NW, KT - 1);
        offB[m] = (int)((L.l21_off + ((size_t)t * 64 + lane) * NS) * ES);
        offC[m] = (int)((L.cidx_off + (size_t)(t * 4 + kq) * 4) * 2);
#pragma unroll
        for (int v = 0; v < 4; ++v) {  // where row r = PB + 16t + drow lands in the next window
            const int r = PB + 16 * (w + m * NW) + drow<T>(kq, v);
            wdst[m][v] = (r < RW ? r - PB : RWP - 1) * RS + jc * ES;
        }
    }
    const int fdst = fr < PB ? (RW - PB + fr) * RS + fj * ES : (RWP - 1) * RS + fj * ES;
    // Y1 rows of this lane relative to the panel (stores of waves past NP
    // and of columns past nrhs go out of range)
    int ysto[4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
        ysto[v] = (w < NP && colok) ? (jc * (int)ldb + 16 * w + drow<T>(kq, v)) * ES : kQOob;

    auto issue = [&](int p, Ops &o) {
        const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<T *>(rec) + (size_t)p * L.per_panel, 0, (int)(L.per_panel * ES), kQRsrcWord3);
        const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<unsigned short *>(idxA) + (size_t)p * L.idx_u16, 0, (int)(L.idx_u16 * 2), kQRsrcWord3);
        bload_vec<T, NS>(rr, offA, o.a);
#pragma unroll
        for (int m = 0; m < MT; ++m) bload_vec<T, NS>(rr, offB[m], o.b[m]);
        const trs_v4u gi = __builtin_amdgcn_raw_buffer_load_b128(ri, kq * NS * 2, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o.g[2 * q] = (int)(gi[q] & 0xffffu);
            o.g[2 * q + 1] = (int)(gi[q] >> 16);
        }
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const trs_v2u ci = __builtin_amdgcn_raw_buffer_load_b64(ri, offC[m], 0, 0);
            o.c[m][0] = (int)(ci[0] & 0xffffu);
            o.c[m][1] = (int)(ci[0] >> 16);
            o.c[m][2] = (int)(ci[1] & 0xffffu);
            o.c[m][3] = (int)(ci[1] >> 16);
        }
        const int row = p * PB + RW + fr;
        o.fresh = qbload<T>(rB, (fr < PB && row < n && c0 + fj < nrhs) ? (fj * (int)ldb + row) * ES : kQOob);
    };

    unsigned char *lds = tf_raw;
    // one panel: uses `u`, starts the loads of panel p+2 into `ld`
    auto panel = [&](int p, auto curc, const Ops &u, Ops &ld) {
        constexpr int cur = decltype(curc)::value;
        const int j0 = p * PB;
        const unsigned char *Wc = lds + cur * WB + jc * ES;
        unsigned char *Wn = lds + (cur ^ 1) * WB;
        issue(min(p + 2, np - 1), ld);
        // the gathered rows (perm W)[0:PB] as B operands
        T g0[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) g0[s] = *reinterpret_cast<const T *>(Wc + u.g[s]);
        // W' = (perm W)[PB:RW] + (-M21) (perm W)[0:PB]
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int t = w + m * NW;
            if (t < KT) {
                v4 z;
#pragma unroll
                for (int v = 0; v < 4; ++v) z[v] = *reinterpret_cast<const T *>(Wc + u.c[m][v]);
#pragma unroll
                for (int s = 0; s < NS; ++s) z = Mf<T>::mma(u.b[m][s], g0[s], z);
#pragma unroll
                for (int v = 0; v < 4; ++v) *reinterpret_cast<T *>(Wn + wdst[m][v]) = z[v];
            }
        }
        *reinterpret_cast<T *>(Wn + fdst) = u.fresh;
        // Y1 = L11^{-1} (perm W)[0:PB], to B (rows past n, columns past
        // nrhs and waves past NP: dropped)
        v4 y = {T(0), T(0), T(0), T(0)};
        if (w < NP) {
#pragma unroll
            for (int s = 0; s < NS; ++s) y = Mf<T>::mma(u.a[s], g0[s], y);
        }
        const bool full = j0 + PB <= n;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * w + drow<T>(kq, v);
            const int off = (full || j0 + r < n) && ysto[v] != kQOob ? ysto[v] + j0 * ES : kQOob;
            qbstore<T>(rB, off, y[v]);
        }
        lds_barrier();
    };

    Ops S0, S1, S2;
    issue(0, S0);
    issue(min(1, np - 1), S1);
    // window of panel 0 (buffer 0)
    for (int e = tid; e < (RWP - 1) * 16; e += 64 * NW) {
        const int r = e >> 4, j = e & 15;
        *reinterpret_cast<T *>(lds + r * RS + j * ES) =
            qbload<T>(rB, (r < RW && r < n && c0 + j < nrhs) ? (j * (int)ldb + r) * ES : kQOob);
    }
    lds_barrier();
    // six panels per trip (three operand sets x two window buffers); panels
    // past the last store nothing: their rows are past n
    const std::integral_constant<int, 0> b0;
    const std::integral_constant<int, 1> b1;
    for (int p = 0; p < np; p += 6) {
        panel(p, b0, S0, S2);
        panel(p + 1, b1, S1, S0);
        panel(p + 2, b0, S2, S1);
        panel(p + 3, b1, S0, S2);
        panel(p + 4, b0, S1, S0);
        panel(p + 5, b1, S2, S1);
    }
}

// ---------------------------------------------------------------- backward sweep
// B <- U^{-1} B (U: upper band of width kv = kl+ku) for 16 columns; the U
// operands come from the pre-pass records.
template <typename T, int PB, int NW, int MT>
__global__ __launch_bounds__(64 * NW) void trs_bwd(T *B, i64 ldb, int nrhs, TrsLayout<T, PB> L, const T *rec, int RB) {
    using v4 = typename Mf<T>::v4;
    constexpr int SW = TStride<T>::value;
    constexpr int NS = PB / 4;
    constexpr int NP = PB / 16;  // (MT: V tiles per wave)
    constexpr int ES = (int)sizeof(T);
    constexpr int RS = SW * ES;
    using Ops = PanelOps<T, NS, MT>;
    extern __shared__ __align__(16) unsigned char tb_raw[];
    const int n = L.n, KTU = L.KTU, np = L.np;
    const int top = 16 * KTU;  // rows above a panel that it updates (>= kv)
    unsigned char *Yr = tb_raw;  // ring [RB][SW], row i in slot i & (RB-1)
    const int msk = RB - 1;
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int jc = lane & 15, kq = lane >> 4;
    const int fr = tid >> 4, fj = tid & 15;
    const int c0 = 16 * blockIdx.x;
    const bool colok = c0 + jc < nrhs;
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(B + (i64)c0 * ldb, 0, (int)(16 * ldb * ES),
                                                                        kQRsrcWord3);
    const int offA = (int)((L.uinv_off + ((size_t)min(w, NP - 1) * 64 + lane) * NS) * ES);
    int offB[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
        offB[m] = (int)((L.ublk_off + ((size_t)min(w + m * NW, KTU - 1) * 64 + lane) * NS) * ES);
    // lane parts of the ring addresses: B operand rows 4s+kq, result rows drow(kq, v)
    const int yb = kq * RS + jc * ES;
    int zb[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) zb[v] = drow<T>(kq, v) * RS + jc * ES;
    int xsto[4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
        xsto[v] = (w < NP && colok) ? (jc * (int)ldb + 16 * w + drow<T>(kq, v)) * ES : kQOob;

    auto issue = [&](int p, Ops &o) {
        const int j0 = p * PB;
        const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<T *>(rec) + (size_t)p * L.per_panel, 0, (int)(L.per_panel * ES), kQRsrcWord3);
        bload_vec<T, NS>(rr, offA, o.a);
#pragma unroll
        for (int m = 0; m < MT; ++m) bload_vec<T, NS>(rr, offB[m], o.b[m]);
        const int frow = j0 - top - PB + fr;
        o.fresh = qbload<T>(rB, (fr < PB && frow >= 0 && c0 + fj < nrhs) ? (fj * (int)ldb + frow) * ES : kQOob);
    };

    // panel index q counts from the last panel (q = 0 is panel np-1).  Ring
    // slots of negative rows (first panels) take throwaway updates: they are
    // never read, as RB >= top + 2 PB keeps them clear of live rows.
    auto panel = [&](int q, const Ops &u, Ops &ld) {
        const int p = np - 1 - q;
        const int j0 = p * PB;
        issue(max(p - 2, 0), ld);
        // Y[j0:j0+PB] as B operands (j0 & msk is a multiple of PB: no wrap inside)
        const unsigned char *yrow = Yr + (j0 & msk) * RS + yb;
        T y0[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) y0[s] = *reinterpret_cast<const T *>(yrow + 4 * s * RS);
        // Y[j0-top : j0] += (-V) Y[j0:j0+PB]
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int t = w + m * NW;
            if (t < KTU) {
                unsigned char *zrow = Yr + ((j0 - top + 16 * t) & msk) * RS;
                v4 z;
#pragma unroll
                for (int v = 0; v < 4; ++v) z[v] = *reinterpret_cast<const T *>(zrow + zb[v]);
#pragma unroll
                for (int s = 0; s < NS; ++s) z = Mf<T>::mma(u.b[m][s], y0[s], z);
#pragma unroll
                for (int v = 0; v < 4; ++v) *reinterpret_cast<T *>(zrow + zb[v]) = z[v];
            }
        }
        const int frow = j0 - top - PB + fr;
        if (fr < PB && frow >= 0) *reinterpret_cast<T *>(Yr + (frow & msk) * RS + fj * ES) = u.fresh;
        // X1 = U11^{-1} Y[j0:j0+PB], to B
        v4 x = {T(0), T(0), T(0), T(0)};
        if (w < NP) {
#pragma unroll
            for (int s = 0; s < NS; ++s) x = Mf<T>::mma(u.a[s], y0[s], x);
        }
        const bool full = j0 >= 0 && j0 + PB <= n;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * w + drow<T>(kq, v);
            const int off = (full || (j0 + r >= 0 && j0 + r < n)) && xsto[v] != kQOob ? xsto[v] + j0 * ES : kQOob;
            qbstore<T>(rB, off, x[v]);
        }
        lds_barrier();
    };

    Ops S0, S1, S2;
    issue(np - 1, S0);
    issue(max(np - 2, 0), S1);
    // ring: rows [jl - top, jl + PB) of the last panel (zero past n)
    const int jl = (np - 1) * PB;
    for (int e = tid; e < (top + PB) * 16; e += 64 * NW) {
        const int rr = e >> 4, j = e & 15;
        const int row = jl - top + rr;
        *reinterpret_cast<T *>(Yr + (row & msk) * RS + j * ES) =
            qbload<T>(rB, (row >= 0 && row < n && c0 + j < nrhs) ? (j * (int)ldb + row) * ES : kQOob);
    }
    lds_barrier();
    // panels before the first (at most two) touch only negative rows: their
    // stores are dropped and their ring updates land in throwaway slots
    for (int q = 0; q < np; q += 3) {
        panel(q, S0, S2);
        panel(q + 1, S1, S0);
        panel(q + 2, S2, S1);
    }
}

constexpr int kTrsPB = 32;
constexpr int kTrsNW = 8;

}  // namespace

// Bytes of the pre-pass output for (n, kl, ku): 0 when the blocked solve does
// not apply (then the caller uses bh_?gbtrs).
template <typename T>
size_t trs_blocked_bytes(i64 n, i64 kl, i64 ku) {
    using Lt = TrsLayout<T, kTrsPB>;
    constexpr int SW = TStride<T>::value;
    // 16 columns of B (leading dimension >= n) must fit one 31-bit buffer offset range
    if (n <= 0 || kl < 1 || 16 * n * (i64)sizeof(T) >= kQOob) return 0;
    // update tiles per wave that the registers hold without spilling:
    // FP32 3 forward, 4 backward; FP64 1 forward, 2 backward
    const i64 mtf = sizeof(T) == 4 ? 3 : 1, mtb = sizeof(T) == 4 ? 4 : 2;
    if (kl > 16 * mtf * kTrsNW || kl + ku > 16 * mtb * kTrsNW) return 0;
    const Lt L = Lt::make(n, kl, ku);
    const int RWP = 16 * ((L.RW + 15) / 16) + 1;
    const size_t fwd_lds = (size_t)(2 * RWP) * SW * sizeof(T);
    int RB = 1;
    while (RB < 16 * L.KTU + 2 * kTrsPB) RB <<= 1;
    const size_t bwd_lds = (size_t)RB * SW * sizeof(T);
    const size_t prep_lds = ((size_t)(L.RW + 2 * kTrsPB + 16 * L.KTU) * (kTrsPB + 1)) * sizeof(T) + 2 * L.RW * sizeof(int);
    if (fwd_lds > 160 * 1024 || bwd_lds > 160 * 1024 || prep_lds > 160 * 1024) return 0;
    const size_t recs = (size_t)L.np * L.per_panel * sizeof(T);
    const size_t idxs = (size_t)L.np * L.idx_u16 * 2;
    if (L.per_panel * sizeof(T) >= (size_t)kQOob) return 0;
    return ((recs + 255) & ~(size_t)255) + ((idxs + 255) & ~(size_t)255);
}

template <typename T>
int trs_blocked_prepare(i64 n, i64 kl, i64 ku, const T *ab, i64 ldab, const i64 *ipiv, void *work, hipStream_t st) {
    using Lt = TrsLayout<T, kTrsPB>;
    const Lt L = Lt::make(n, kl, ku);
    const size_t recs = (size_t)L.np * L.per_panel * sizeof(T);
    T *rec = (T *)work;
    unsigned short *idx = (unsigned short *)((char *)work + ((recs + 255) & ~(size_t)255));
    const size_t lds = ((size_t)(L.RW + 2 * kTrsPB + 16 * L.KTU) * (kTrsPB + 1)) * sizeof(T) + 2 * L.RW * sizeof(int);
    auto kern = trs_prep<T, kTrsPB>;
    if (lds > 64 * 1024)
        BH_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)L.np), dim3(256), lds, st, ab, ldab, ipiv, L, idx, rec);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
int trs_blocked_solve(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ldab, const void *work, T *B, i64 ldb,
                      hipStream_t st) {
    using Lt = TrsLayout<T, kTrsPB>;
    constexpr int SW = TStride<T>::value;
    if (nrhs <= 0 || n <= 0) return BH_OK;
    if (16 * ldb * (i64)sizeof(T) >= kQOob) {
        set_error("gbtrs_blocked: ldb = %lld is too large for 16 columns in one buffer range", (long long)ldb);
        return BH_ERR_WORKSPACE;
    }
    const Lt L = Lt::make(n, kl, ku);
    const size_t recs = (size_t)L.np * L.per_panel * sizeof(T);
    const T *rec = (const T *)work;
    const unsigned short *idx = (const unsigned short *)((const char *)work + ((recs + 255) & ~(size_t)255));
    const dim3 grid((unsigned)((nrhs + 15) / 16));
    const int RWP = 16 * ((L.RW + 15) / 16) + 1;
    const size_t fl = (size_t)(2 * RWP) * SW * sizeof(T);
    const int mtf = (L.KT + kTrsNW - 1) / kTrsNW;
    void (*kf)(T *, i64, int, Lt, const unsigned short *, const T *) = nullptr;
    if constexpr (sizeof(T) == 4) {
        if (mtf > 3) return BH_ERR_WORKSPACE;
        kf = mtf <= 1 ? trs_fwd<T, kTrsPB, kTrsNW, 1>
           : mtf == 2 ? trs_fwd<T, kTrsPB, kTrsNW, 2>
                      : trs_fwd<T, kTrsPB, kTrsNW, 3>;
    } else {
        if (mtf > 1) return BH_ERR_WORKSPACE;
        kf = trs_fwd<T, kTrsPB, kTrsNW, 1>;
    }
    if (fl > 64 * 1024) BH_HIP(hipFuncSetAttribute((const void *)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fl));
    hipLaunchKernelGGL(kf, grid, dim3(64 * kTrsNW), fl, st, B, ldb, (int)nrhs, L, idx, rec);
    BH_LAUNCH_CHECK();
    int RB = 1;
    while (RB < 16 * L.KTU + 2 * kTrsPB) RB <<= 1;
    const size_t bl = (size_t)RB * SW * sizeof(T);
    const int mtb = (L.KTU + kTrsNW - 1) / kTrsNW;
    void (*kb)(T *, i64, int, Lt, const T *, int) = nullptr;
    if constexpr (sizeof(T) == 4) {
        kb = mtb <= 1 ? trs_bwd<T, kTrsPB, kTrsNW, 1>
           : mtb == 2 ? trs_bwd<T, kTrsPB, kTrsNW, 2>
           : mtb == 3 ? trs_bwd<T, kTrsPB, kTrsNW, 3>
                      : trs_bwd<T, kTrsPB, kTrsNW, 4>;
    } else {
        if (mtb > 2) return BH_ERR_WORKSPACE;
        kb = mtb <= 1 ? trs_bwd<T, kTrsPB, kTrsNW, 1> : trs_bwd<T, kTrsPB, kTrsNW, 2>;
    }
    if (bl > 64 * 1024) BH_HIP(hipFuncSetAttribute((const void *)kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bl));
    hipLaunchKernelGGL(kb, grid, dim3(64 * kTrsNW), bl, st, B, ldb, (int)nrhs, L, rec, RB);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template size_t trs_blocked_bytes<float>(i64, i64, i64);
template size_t trs_blocked_bytes<double>(i64, i64, i64);
template int trs_blocked_prepare<float>(i64, i64, i64, const float *, i64, const i64 *, void *, hipStream_t);
template int trs_blocked_prepare<double>(i64, i64, i64, const double *, i64, const i64 *, void *, hipStream_t);
template int trs_blocked_solve<float>(i64, i64, i64, i64, const float *, i64, const void *, float *, i64, hipStream_t);
template int trs_blocked_solve<double>(i64, i64, i64, i64, const double *, i64, const void *, double *, i64,
                                       hipStream_t);

}  // namespace bh

using namespace bh;

template <typename T>
static int gbtrs_blocked_impl(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ldab, const i64 *ipiv, T *B, i64 ldb,
                              void *work, size_t lwork) {
    BH_ARG(n >= 0, 1);
    BH_ARG(kl >= 0, 2);
    BH_ARG(ku >= 0, 3);
    BH_ARG(nrhs >= 0, 4);
    BH_ARG(ldab >= 2 * kl + ku + 1, 6);
    BH_ARG(ldb >= imax(1, n), 9);
    if (n == 0 || nrhs == 0) return BH_OK;
    const size_t need = trs_blocked_bytes<T>(n, kl, ku);
    if (need == 0) {
        set_error("gbtrs_blocked: bands (%lld,%lld) are outside the blocked solve's range (FP32: 1 <= kl <= 384, "
                  "kl+ku <= 512; FP64: 1 <= kl <= 128, kl+ku <= 256); use bh_?gbtrs", (long long)kl, (long long)ku);
        return BH_ERR_WORKSPACE;
    }
    if (work == nullptr || lwork < need) {
        set_error("gbtrs_blocked: workspace of %zu bytes needed, %zu given", need, lwork);
        return BH_ERR_WORKSPACE;
    }
    int st = trs_blocked_prepare<T>(n, kl, ku, ab, ldab, ipiv, work, stream());
    if (st) return st;
    return trs_blocked_solve<T>(n, kl, ku, nrhs, ab, ldab, work, B, ldb, stream());
}

extern "C" {

int bh_gbtrs_blocked_workspace(char prec, bh_int n, bh_int kl, bh_int ku, size_t *bytes) {
    BH_ARG(prec == 'd' || prec == 's', 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG(bytes != nullptr, 5);
    *bytes = prec == 'd' ? trs_blocked_bytes<double>(n, kl, ku) : trs_blocked_bytes<float>(n, kl, ku);
    return BH_OK;
}

int bh_dgbtrs_blocked(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *AB, bh_int ldab, const bh_int *ipiv,
                      double *B, bh_int ldb, void *work, size_t lwork) {
    return gbtrs_blocked_impl<double>(n, kl, ku, nrhs, AB, ldab, ipiv, B, ldb, work, lwork);
}

int bh_sgbtrs_blocked(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const float *AB, bh_int ldab, const bh_int *ipiv,
                      float *B, bh_int ldb, void *work, size_t lwork) {
    return gbtrs_blocked_impl<float>(n, kl, ku, nrhs, AB, ldab, ipiv, B, ldb, work, lwork);
}

}  // extern "C"
