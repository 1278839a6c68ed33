// bh_lu_small.hip -- band LU (and fused solve) with one wave per system, for
// bands whose window has at most 64 columns (kl + ku <= 63).
//
// Arithmetic of LAPACK xGBTF2 (LAPACK.gbtrf!, src/banded/BandedLU.jl:98) and
// xGBTRS 'N' (ldiv!, src/banded/linalg.jl:24-30), bit for bit: IDAMAX pivot
// (first largest |a|), interchange over the live columns, multipliers
// a * (1/pivot), updates fma(l, -u, a) on rows j+1..j+km and columns j+1..ju;
// forward solve fma(l, -b_j, b_i) after the interchange, back substitution
// x_j = b_j / u_jj, fma(u_ij, -x_j, b_i).
//
// Layout: lane c mod 64 holds band column c of the window (columns j..j+kl+ku)
// with its rows j..j+kl in registers y[0..kl].  Step j: the pivot column is
// broadcast (readlane), every lane finds the pivot and the multipliers
// (uniform), swaps its rows 0 and jp, and the lanes of columns j+1..ju
// update their rows; row j leaves (U row), rows shift up through the
// registers, the row entering at the bottom is prefetched.  The lane of the
// retired column j takes, when the window reaches it, column j+64.
//
// Used for the single-system gbtrf of narrow bands and, with the solve fused
// into the factorisation's sweep, for the batched gbsv (many independent
// systems in one persistent launch; each wave walks its share of the
// schedule).
#include "bh_common.hpp"
#include "bh_lu_lane.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <type_traits>

namespace bh {

namespace {

constexpr int SPF = 8;  // prefetch distance (steps)

template <int K>
struct SK {
    static constexpr int value = K;
};
template <int N>
struct SUnroll {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        SUnroll<N - 1>::run(f);
        f(SK<N - 1>{});
    }
};
template <>
struct SUnroll<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};

__device__ __forceinline__ int smin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int smax(int a, int b) { return a > b ? a : b; }

// Factor one square-or-rectangular system in place (AB padded, ldab >= 2kl+ku+1);
// with SOLVE (m == n) also overwrite b with A \ b.  Returns LAPACK info.
// Every lane of the wave calls it with the same arguments.
template <typename T, int RL, bool SOLVE>
__device__ int small_gbsv_wave(int m, int n, int kl, int ku, T *__restrict__ ab, i64 ld, i64 *__restrict__ ipiv,
                               T *__restrict__ b) {
    const int lane = (int)__lane_id();
    const int kv = kl + ku;
    const int mn = smin(m, n);
    if (mn == 0) return 0;
    auto orig = [&](int r, int c) -> T {
        if (r < 0 || r >= m || c < 0 || c >= n || c - r > ku || r - c > kl) return T(0);
        return ab[(i64)c * ld + (kv + r - c)];
    };
    auto col_of = [&](int j) { return j + ((lane - j) & 63); };

    T y[RL];
#pragma unroll
    for (int i = 0; i < RL; ++i) y[i] = (lane <= kv && i <= kl) ? orig(i, lane) : T(0);
    T rowin[SPF];
    auto fetch_row = [&](int slot, int j) {
        const int c = col_of(j + 1);
        rowin[slot] = (j < mn && c - (j + 1) <= kv) ? orig(j + 1 + kl, c) : T(0);
    };
    T bw[RL], bin[SPF];
    auto fetch_b = [&](int slot, int j) {
        const int e = j + 1 + kl;
        bin[slot] = (SOLVE && j < mn && e < n) ? b[e] : T(0);
    };
    if (SOLVE) {
#pragma unroll
        for (int i = 0; i < RL; ++i) bw[i] = (i <= kl && i < n) ? b[i] : T(0);
    }
    SUnroll<SPF>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        fetch_row(s, s);
        if (SOLVE) fetch_b(s, s);
    });

    int ju = 0, info = 0;
    for (int j0 = 0; j0 < mn; j0 += SPF) {
        SUnroll<SPF>::run([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            const int j = j0 + s;
            if (j < mn) {
                const int own = j & 63;
                const int km = smin(kl, m - 1 - j);
                T x[RL];
#pragma unroll
                for (int i = 0; i < RL; ++i) x[i] = lane_read(y[i], own);
                // pivot: first largest |x_i|, i <= km (NaN loses to every number)
                T best = dabs(x[0]);
                if (!(best == best)) best = T(-0.5);
                int jp = 0;
#pragma unroll
                for (int i = 1; i < RL; ++i) {
                    if (i <= km) {
                        T v = dabs(x[i]);
                        if (!(v == v)) v = T(-0.5);
                        if (v > best) {
                            best = v;
                            jp = i;
                        }
                    }
                }
                T pv = x[0];
#pragma unroll
                for (int i = 1; i < RL; ++i)
                    if (i == jp) pv = x[i];
                const bool nz = (pv != T(0));
                if (nz) {
                    ju = smax(ju, smin(j + jp + ku, n - 1));
                    if (jp != 0) {
#pragma unroll
                        for (int i = 1; i < RL; ++i) {
                            if (i == jp) {
                                const T t0 = y[0];
                                y[0] = y[i];
                                y[i] = t0;
                                x[i] = x[0];
                                x[0] = pv;
                            }
                        }
                    }
                } else if (info == 0) {
                    info = j + 1;
                }
                const T rcp = T(1) / pv;
                T lm[RL];
#pragma unroll
                for (int i = 1; i < RL; ++i) lm[i] = (i <= km) ? (nz ? x[i] * rcp : x[i]) : T(0);
                const int c = col_of(j);
                if (nz && c > j && c <= ju) {
                    const T nu = -y[0];
#pragma unroll
                    for (int i = 1; i < RL; ++i)
                        if (i <= km) y[i] = dfma(lm[i], nu, y[i]);
                }
                // U row j (y[0] of the window columns), column j's multipliers, pivot
                if (lane == own) {
                    T *colj = ab + (i64)j * ld + kv;
                    colj[0] = x[0];
#pragma unroll
                    for (int i = 1; i < RL; ++i)
                        if (i <= km) colj[i] = lm[i];
                    ipiv[j] = (i64)(j + jp + 1);
                } else if (c > j && c - j <= kv && c < n) {
                    ab[(i64)c * ld + (kv - (c - j))] = y[0];
                }
                if (SOLVE) {
                    // forward substitution step j (xGBTRS: swap, then the L column)
                    if (jp != 0) {
#pragma unroll
                        for (int i = 1; i < RL; ++i) {
                            if (i == jp) {
                                const T t0 = bw[0];
                                bw[0] = bw[i];
                                bw[i] = t0;
                            }
                        }
                    }
                    const T nb = -bw[0];
#pragma unroll
                    for (int i = 1; i < RL; ++i)
                        if (i <= smin(kl, n - 1 - j)) bw[i] = dfma(lm[i], nb, bw[i]);
                    if (lane == 0) b[j] = bw[0];
#pragma unroll
                    for (int i = 0; i + 1 < RL; ++i) bw[i] = bw[i + 1];
                    bw[RL - 1] = T(0);
#pragma unroll
                    for (int i = 0; i < RL; ++i)
                        if (i == kl) bw[i] = bin[s];
                    fetch_b(s, j + SPF);
                }
                // retire column j (its lane holds nothing until column j+64 enters)
                if (lane == own) {
#pragma unroll
                    for (int i = 0; i < RL; ++i) y[i] = T(0);
                }
#pragma unroll
                for (int i = 0; i + 1 < RL; ++i) y[i] = y[i + 1];
                y[RL - 1] = T(0);
#pragma unroll
                for (int i = 0; i < RL; ++i)
                    if (i == kl) y[i] = rowin[s];
                fetch_row(s, j + SPF);
            }
        });
    }
    // Fill-in rows LAPACK zeroes that no U row covered (rows >= m).
    {
        const int c_lo = smax(ku + 1, m + ku), c_hi = smin(n, mn + kv);
        for (int c = c_lo + lane; c < c_hi; c += 64) {
            for (int q = smax(0, kv - c); q < kl; ++q)
                if (c - kv + q >= m) ab[(i64)c * ld + q] = T(0);
        }
    }
    if (SOLVE) {
        // back substitution: lane t holds b_{j-t}, t = 0 .. kl+ku
        T w = (lane <= kv && n - 1 - lane >= 0) ? b[n - 1 - lane] : T(0);
        T ur[SPF], br[SPF];
        auto fetch_u = [&](int slot, int j) {
            ur[slot] = (j >= 0 && lane <= smin(kv, j)) ? ab[(i64)j * ld + kv - lane] : T(0);
            const int e = j - 1 - kv;
            br[slot] = (j >= 0 && e >= 0) ? b[e] : T(0);
        };
        SUnroll<SPF>::run([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            fetch_u(s, n - 1 - s);
        });
        for (int j0 = n - 1; j0 >= 0; j0 -= SPF) {
            SUnroll<SPF>::run([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                const int j = j0 - s;
                if (j >= 0) {
                    const T xj = lane_read(w, 0) / lane_read(ur[s], 0);
                    if (lane == 0) b[j] = xj;
                    if (lane >= 1 && lane <= smin(kv, j)) w = dfma(ur[s], -xj, w);
                    w = wave_down1(w, T(0));
                    if (lane == kv) w = br[s];
                    fetch_u(s, j - SPF);
                }
            });
        }
    }
    return info;
}

// ---------------------------------------------------------------- single system
template <typename T, int RL>
__global__ void __launch_bounds__(64) gbtrf_small_kernel(int m, int n, int kl, int ku, T *ab, i64 ld, i64 *ipiv,
                                                         i64 *info) {
    const int r = small_gbsv_wave<T, RL, false>(m, n, kl, ku, ab, ld, ipiv, nullptr);
    if (threadIdx.x == 0) *info = r;
}

// ---------------------------------------------------------------- batched
// (the batched kernel and its descriptor live in bh_lu_lane.hpp)

// Register-row classes: kl + 1 <= RL for RL in {2, 3, 5, 9, 12, 17, 24, 32}.
constexpr int kSmallRL[] = {2, 3, 5, 9, 12, 17, 24, 32};
int small_class(i64 kl) {
    for (int c = 0; c < 8; ++c)
        if (kl + 1 <= kSmallRL[c]) return c;
    return -1;
}

}  // namespace

// Single-system narrow-band LU (kl + ku <= 63, kl <= 31), or -1.  The
// rotating-register kernel runs for kl <= 8 and kl = 16 when the band
// storage is under 2 GiB (its buffer offsets are 32-bit); other widths, and
// BANDEDHIP_SMALL_LU=class, take the register-class kernel.
template <typename T>
int small_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st) {
    if (kl + ku > 63 || m >= (1ll << 30) || n >= (1ll << 30)) return -1;
    const int cls = small_class(kl);
    if (cls < 0) return -1;
    const int mi = (int)m, ni = (int)n, li = (int)kl, ui = (int)ku;
    const char *e = getenv("BANDEDHIP_SMALL_LU");
    if (!(e && strcmp(e, "class") == 0)) {
        const int r = lane_gbtrf_launch(m, n, kl, ku, ab, ldab, ipiv, dinfo, st);
        if (r != -1) return r;
    }
#define BH_SMALL_CASE(C, RLv)                                                                                  \
    case C:                                                                                                    \
        hipLaunchKernelGGL((gbtrf_small_kernel<T, RLv>), dim3(1), dim3(64), 0, st, mi, ni, li, ui, ab, ldab, ipiv, \
                           dinfo);                                                                             \
        break;
    switch (cls) {
        BH_SMALL_CASE(0, 2)
        BH_SMALL_CASE(1, 3)
        BH_SMALL_CASE(2, 5)
        BH_SMALL_CASE(3, 9)
        BH_SMALL_CASE(4, 12)
        BH_SMALL_CASE(5, 17)
        BH_SMALL_CASE(6, 24)
        BH_SMALL_CASE(7, 32)
    }
#undef BH_SMALL_CASE
    BH_LAUNCH_CHECK();
    return BH_OK;
}
template int small_gbtrf_launch<double>(i64, i64, i64, i64, double *, i64, i64 *, i64 *, hipStream_t);
template int small_gbtrf_launch<float>(i64, i64, i64, i64, float *, i64, i64 *, i64 *, hipStream_t);

}  // namespace bh

// ====================================================================== C ABI
using namespace bh;

struct bh_batch_plan {
    i64 nsys = 0;
    std::vector<i64> n, kl, ku;  // caller's order
    std::vector<i64> ab_off, vec_off;
    i64 ab_elems = 0, vec_elems = 0;
    BatchDesc *desc = nullptr;  // device, schedule order, padded to whole rounds
    i64 desc_len = 0, desc_cap = 0;
    int waves_per_block = 1, blocks = 0, cus = 256, order_kind = 0;
    // graphs captured from this plan: each baked in desc, desc_len and the launch
    // geometry, so the schedule may not change (nor the plan go) while any exists
    int live_graphs = 0;
};

struct bh_batch_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bh_batch_plan *plan = nullptr;
};

static int plan_busy(const bh_batch_plan *p, const char *what) {
    if (p->live_graphs == 0) return BH_OK;
    bh::set_error("%s: the plan has %d captured graph(s); destroy them first", what, p->live_graphs);
    return BH_ERR_BUSY;
}

namespace {

// Schedule: W = blocks * waves_per_block resident waves; wave w solves the
// systems at positions w, w + W, ... of desc.  order_kind 0: decreasing work
// dealt out in snake rounds (round r goes to waves 0..W-1, round r+1 back
// from W-1), 1: the caller's order round-robin, 2: the band classes
// interleaved, decreasing work within each class, snake rounds.
int plan_set_order(bh_batch_plan *p, int kind) {
    const i64 ns = p->nsys;
    const i64 W = (i64)p->blocks * p->waves_per_block;
    std::vector<int> order(ns);
    for (i64 s = 0; s < ns; ++s) order[s] = (int)s;
    auto work = [&](int s) { return (double)p->n[s] * (p->kl[s] + 1) * (p->kl[s] + p->ku[s] + 1); };
    if (kind == 0) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return work(a) > work(c); });
    } else if (kind == 2) {
        std::vector<std::vector<int>> by(8);
        for (i64 s = 0; s < ns; ++s) by[small_class(p->kl[s])].push_back((int)s);
        for (auto &v : by) std::stable_sort(v.begin(), v.end(), [&](int a, int c) { return work(a) > work(c); });
        order.clear();
        for (size_t k = 0; order.size() < (size_t)ns; ++k)
            for (auto &v : by)
                if (k < v.size()) order.push_back(v[k]);
    }
    const i64 rounds = (ns + W - 1) / W;
    const i64 len = std::max<i64>(rounds * W, 1);
    std::vector<BatchDesc> d(len, BatchDesc{0, 0, 0, 0, 0, 0, -1});
    for (i64 i = 0; i < ns; ++i) {
        const i64 r = i / W, k = i % W;
        const i64 w = (kind == 1 || (r & 1) == 0) ? k : W - 1 - k;
        const int s = order[i];
        d[r * W + w] = BatchDesc{p->ab_off[s], p->vec_off[s], (int)p->n[s], (int)p->kl[s], (int)p->ku[s],
                                 small_class(p->kl[s]), s};
    }
    if (len > p->desc_cap) {
        (void)hipFree(p->desc);
        p->desc = nullptr;
        p->desc_cap = 0;
        if (hipMalloc((void **)&p->desc, sizeof(BatchDesc) * len) != hipSuccess) {
            set_error("bh_batch plan: out of device memory");
            return BH_ERR_ALLOC;
        }
        p->desc_cap = len;
    }
    BH_HIP(hipMemcpy(p->desc, d.data(), sizeof(BatchDesc) * len, hipMemcpyHostToDevice));
    p->desc_len = len;
    p->order_kind = kind;
    return BH_OK;
}

template <typename T>
int batch_launch(bh_batch_plan *p, T *AB, T *B, i64 *ipiv, i64 *info, hipStream_t st) {
    if (p->nsys == 0) return BH_OK;
    const int r = lane_batch_launch((unsigned)p->blocks, (unsigned)(64 * p->waves_per_block), st, p->desc,
                                    (int)p->desc_len, AB, B, ipiv, info);
    if (r != BH_OK) return r;
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
int batch_check(bh_batch_plan *p, T *AB, T *B, i64 *ipiv, i64 *info) {
    BH_ARG(p != nullptr, 1);
    BH_ARG(p->nsys == 0 || (AB && B && ipiv && info), 2);
    return BH_OK;
}

template <typename T>
int batch_autotune(bh_batch_plan *p, const T *AB, const T *B, double *best_ms) {
    BH_ARG(p != nullptr, 1);
    BH_ARG(p->nsys == 0 || (AB && B), 2);
    if (p->nsys == 0) return BH_OK;
    // trials run on scratch copies; the caller's arrays are left untouched
    T *ab2 = nullptr, *b2 = nullptr;
    i64 *ip2 = nullptr, *inf2 = nullptr;
    hipStream_t s = stream();
    auto cleanup = [&] {
        (void)hipFree(ab2);
        (void)hipFree(b2);
        (void)hipFree(ip2);
        (void)hipFree(inf2);
    };
    if (hipMalloc((void **)&ab2, sizeof(T) * p->ab_elems) != hipSuccess ||
        hipMalloc((void **)&b2, sizeof(T) * p->vec_elems) != hipSuccess ||
        hipMalloc((void **)&ip2, sizeof(i64) * p->vec_elems) != hipSuccess ||
        hipMalloc((void **)&inf2, sizeof(i64) * p->nsys) != hipSuccess) {
        cleanup();
        set_error("bh_batch_autotune: out of device memory");
        return BH_ERR_ALLOC;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int wpb[] = {1, 2, 4};
    const int mult[] = {1, 2, 4, 8};
    const int kinds[] = {0, 2};
    double best = 1e300;
    int bw = p->waves_per_block, bb = p->blocks, bk = p->order_kind;
    for (int kind : kinds) {
        for (int w : wpb) {
            for (int mu : mult) {
                p->waves_per_block = w;
                p->blocks = std::max(1, p->cus * mu / w);
                if (plan_set_order(p, kind) != BH_OK) continue;
                float ms_sum = 0.f;
                for (int rep = 0; rep < 2; ++rep) {
                    (void)hipMemcpyAsync(ab2, AB, sizeof(T) * p->ab_elems, hipMemcpyDeviceToDevice, s);
                    (void)hipMemcpyAsync(b2, B, sizeof(T) * p->vec_elems, hipMemcpyDeviceToDevice, s);
                    (void)hipEventRecord(e0, s);
                    batch_launch<T>(p, ab2, b2, ip2, inf2, s);
                    (void)hipEventRecord(e1, s);
                    (void)hipEventSynchronize(e1);
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (rep > 0) ms_sum += ms;
                }
                if (ms_sum < best) {
                    best = ms_sum;
                    bw = w;
                    bb = p->blocks;
                    bk = kind;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cleanup();
    p->waves_per_block = bw;
    p->blocks = bb;
    const int so = plan_set_order(p, bk);
    if (so != BH_OK) return so;
    if (best_ms) *best_ms = best;
    BH_HIP(hipGetLastError());
    return BH_OK;
}

template <typename T>
int batch_graph_create(bh_batch_plan *p, T *AB, T *B, i64 *ipiv, i64 *info, bh_batch_graph **out) {
    int st = batch_check<T>(p, AB, B, ipiv, info);
    if (st != BH_OK) return st;
    BH_ARG(out != nullptr, 6);
    auto *g = new bh_batch_graph();
    hipStream_t s = stream();
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        st = batch_launch<T>(p, AB, B, ipiv, info, s);
        hipError_t e2 = hipStreamEndCapture(s, &g->graph);
        if (st == BH_OK && e2 != hipSuccess) st = hip_fail(e2, "hipStreamEndCapture", __FILE__, __LINE__);
    } else {
        st = hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    }
    if (st == BH_OK) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) st = hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    if (st != BH_OK) {
        if (g->exec) hipGraphExecDestroy(g->exec);
        if (g->graph) hipGraphDestroy(g->graph);
        delete g;
        return st;
    }
    g->plan = p;
    ++p->live_graphs;
    *out = g;
    return BH_OK;
}

}  // namespace

extern "C" {

int bh_batch_plan_create(bh_int nsys, const bh_int *n, const bh_int *kl, const bh_int *ku, bh_batch_plan **out) {
    BH_ARG(nsys >= 0, 1);
    BH_ARG(nsys == 0 || (n && kl && ku), 2);
    BH_ARG(out != nullptr, 5);
    *out = nullptr;
    auto *p = new bh_batch_plan();
    p->nsys = nsys;
    if (nsys > 0) {
        p->n.assign(n, n + nsys);
        p->kl.assign(kl, kl + nsys);
        p->ku.assign(ku, ku + nsys);
    }
    for (i64 s = 0; s < nsys; ++s) {
        if (p->n[s] < 0 || p->n[s] >= (1ll << 30) || p->kl[s] < 0 || p->ku[s] < 0 || p->kl[s] + p->ku[s] > 63 ||
            p->kl[s] > 8 || (2 * p->kl[s] + p->ku[s] + 1) * p->n[s] * 8 >= (1ll << 31)) {
            set_error("bh_batch_plan_create: system %lld (n=%lld kl=%lld ku=%lld) is outside the batched kernel "
                      "(kl <= 8, kl+ku <= 63, band storage under 2 GiB)",
                      (long long)s, (long long)p->n[s], (long long)p->kl[s], (long long)p->ku[s]);
            delete p;
            return -2;
        }
        p->ab_off.push_back(p->ab_elems);
        p->vec_off.push_back(p->vec_elems);
        p->ab_elems += (2 * p->kl[s] + p->ku[s] + 1) * p->n[s];
        p->vec_elems += p->n[s];
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev);
    p->waves_per_block = 1;
    p->blocks = std::max(1, p->cus * 4);
    const int st = plan_set_order(p, 0);
    if (st != BH_OK) {
        (void)hipFree(p->desc);
        delete p;
        return st;
    }
    *out = p;
    return BH_OK;
}

int bh_batch_plan_layout(const bh_batch_plan *p, bh_int *ab_off, bh_int *vec_off, bh_int *ab_elems,
                         bh_int *vec_elems) {
    BH_ARG(p != nullptr, 1);
    for (i64 s = 0; s < p->nsys; ++s) {
        if (ab_off) ab_off[s] = p->ab_off[s];
        if (vec_off) vec_off[s] = p->vec_off[s];
    }
    if (ab_elems) *ab_elems = p->ab_elems;
    if (vec_elems) *vec_elems = p->vec_elems;
    return BH_OK;
}

int bh_batch_plan_schedule(bh_batch_plan *p, int order_kind, int waves_per_block, int blocks) {
    BH_ARG(p != nullptr, 1);
    BH_ARG(order_kind >= 0 && order_kind <= 2, 2);
    BH_ARG(waves_per_block >= 1 && waves_per_block <= 4, 3);
    BH_ARG(blocks >= 1, 4);
    {
        const int b = plan_busy(p, "bh_batch_plan_schedule");
        if (b) return b;
    }
    p->waves_per_block = waves_per_block;
    p->blocks = blocks;
    return plan_set_order(p, order_kind);
}

int bh_batch_plan_get_schedule(const bh_batch_plan *p, int *order_kind, int *waves_per_block, int *blocks) {
    BH_ARG(p != nullptr, 1);
    if (order_kind) *order_kind = p->order_kind;
    if (waves_per_block) *waves_per_block = p->waves_per_block;
    if (blocks) *blocks = p->blocks;
    return BH_OK;
}

int bh_batch_plan_destroy(bh_batch_plan *p) {
    if (!p) return BH_OK;
    {
        const int b = plan_busy(p, "bh_batch_plan_destroy");
        if (b) return b;
    }
    (void)hipFree(p->desc);
    delete p;
    return BH_OK;
}

int bh_dgbsv_batched(bh_batch_plan *p, double *AB, double *B, bh_int *ipiv, bh_int *info) {
    const int st = batch_check<double>(p, AB, B, ipiv, info);
    return st != BH_OK ? st : batch_launch<double>(p, AB, B, ipiv, info, stream());
}
int bh_sgbsv_batched(bh_batch_plan *p, float *AB, float *B, bh_int *ipiv, bh_int *info) {
    const int st = batch_check<float>(p, AB, B, ipiv, info);
    return st != BH_OK ? st : batch_launch<float>(p, AB, B, ipiv, info, stream());
}

int bh_dgbsv_batched_autotune(bh_batch_plan *p, const double *AB, const double *B, double *best_ms) {
    if (p && p->live_graphs) return plan_busy(p, "bh_dgbsv_batched_autotune");
    return batch_autotune<double>(p, AB, B, best_ms);
}
int bh_sgbsv_batched_autotune(bh_batch_plan *p, const float *AB, const float *B, double *best_ms) {
    if (p && p->live_graphs) return plan_busy(p, "bh_sgbsv_batched_autotune");
    return batch_autotune<float>(p, AB, B, best_ms);
}

int bh_dgbsv_batched_graph(bh_batch_plan *p, double *AB, double *B, bh_int *ipiv, bh_int *info,
                           bh_batch_graph **g) {
    return batch_graph_create<double>(p, AB, B, ipiv, info, g);
}
int bh_sgbsv_batched_graph(bh_batch_plan *p, float *AB, float *B, bh_int *ipiv, bh_int *info, bh_batch_graph **g) {
    return batch_graph_create<float>(p, AB, B, ipiv, info, g);
}
int bh_batch_graph_launch(bh_batch_graph *g) {
    BH_ARG(g != nullptr, 1);
    BH_HIP(hipGraphLaunch(g->exec, stream()));
    return BH_OK;
}
int bh_batch_graph_destroy(bh_batch_graph *g) {
    if (!g) return BH_OK;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->plan) --g->plan->live_graphs;
    delete g;
    return BH_OK;
}

}  // extern "C"
