// fednl_kernels.hip -- MI355X (gfx950) kernels and host engine of the FedNL
// round. One FedNL round of dopt's single-node simulation
// (dopt/bin_fednl_local/include/TrainLoop.h:1054-1508 with the client loops of
// LocalWorkerForFedNL_{TopK,TopLEK,RandK,RandSeqK}.h) runs as seven kernels on
// one stream, every client simulated in place:
//
//   k_margin        t = A x, sigma, D = (1-sigma)sigma/m, g = (1-sigma)(-1/m), -log sigma
//   k_hessian_f64   upper tiles of A^T diag(D) A on v_mfma_f64_16x16x4f64, fused with
//                   +lambda, Delta = H - learned, Lk partial sums, and the client gradient
//   k_select_*      index selection (heap scan of TopK/TopLEK, RandK shuffle, RandSeqK window)
//   k_emit          message values, learned-Hessian update, TopLEK random rounding
//   k_master_h      H[idx] += val in client order (deterministic)
//   k_master_vec    gradient / Lk / f averages, norms, trace record
//   k_newton        Cholesky of H + shift I, two triangular solves, x -= dx
//
// Device layout (per problem):
//   A       n x m_pad x d_pad  sample rows b_s * a_s (sign folded as dopt's
//           L2RegulirizeLogisticRegression ctor, logistic_regression.h:26-76), zero padded
//   learned n x S   packed upper (list order of SpecialMathRoutinesForMatrix.h:35-64)
//   delta   n x S   packed upper Delta = H_c(x) - learned_c of the current round
//   H       S       packed upper master estimate
//   pos/val n x k   selected list positions (ascending) and values sent to the master
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "fednl_hip.h"

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace fednl {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static void set_err(const char* what, hipError_t e = hipSuccess)
{
    g_err = what;
    if (e != hipSuccess) {
        g_err += ": ";
        g_err += hipGetErrorString(e);
    }
}

#define HIP_OK(expr)                                                   \
    do {                                                               \
        hipError_t _e = (expr);                                        \
        if (_e != hipSuccess) {                                        \
            set_err(#expr, _e);                                        \
            return FEDNL_E_DEVICE;                                     \
        }                                                              \
    } while (0)

// ---------------------------------------------------------------------------
// constants
// ---------------------------------------------------------------------------
constexpr int kTile = 64;          // Hessian output tile (rows and columns)
constexpr int kKC = 16;            // samples staged per LDS chunk
constexpr int kPadLds = 4;         // LDS row padding (doubles)
constexpr int kSeg = 1024;         // heap-scan staging segment (list positions)
constexpr int kMaxHeapK = 8192;    // LDS-resident heap limit
constexpr int kWin = 1024;         // master scatter window (list positions)
constexpr int kMaxClients = 1024;  // master scatter client table
constexpr int kNB = 32;            // Cholesky panel width
constexpr int kMaxNewtonD = 512;   // single-workgroup Newton solve limit (LDS panel)

__host__ __device__ inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// flat index j*lda + i of list position p (upper triangle, column by column)
__device__ inline uint32_t pos_col(uint64_t p)
{
    uint64_t j = (uint64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (j * (j + 1) / 2 > p) --j;
    while ((j + 1) * (j + 2) / 2 <= p) ++j;
    return (uint32_t)j;
}

__device__ inline uint32_t pos_flat(uint64_t p, int lda)
{
    uint64_t j = pos_col(p);
    uint64_t i = p - j * (j + 1) / 2;
    return (uint32_t)(j * (uint64_t)lda + i);
}

// list position p is a diagonal entry (j, j) iff p = j(j+3)/2, i.e. 8p+9 = (2j+3)^2
__device__ inline bool pos_is_diag(uint64_t p)
{
    uint64_t v = 8 * p + 9;
    uint64_t r = (uint64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r * r == v;
}

// ---------------------------------------------------------------------------
// LCGs of dopt (RandomGenIntegerLinear.h:49-123, RandomGenRealLinear.h:49-54)
// ---------------------------------------------------------------------------
constexpr uint32_t kLcgA = 214013u;
constexpr uint32_t kLcgC = 2531011u;

__device__ inline uint32_t lcg_real_step(uint32_t x) { return (kLcgA * x + kLcgC) % 4294967295u; }

// n-fold composition of x -> a x + c (mod 2^32), as (A, C) with x_n = A x_0 + C
__device__ inline void lcg_jump(uint64_t n, uint32_t& A, uint32_t& C)
{
    uint32_t ra = 1u, rc = 0u, ba = kLcgA, bc = kLcgC;
    while (n) {
        if (n & 1) {
            ra = ba * ra;
            rc = ba * rc + bc;
        }
        bc = ba * bc + bc;
        ba = ba * ba;
        n >>= 1;
    }
    A = ra;
    C = rc;
}

// SimpleMathRoutines.h:407-414 (TInteger = size_t)
__device__ inline uint64_t round_nearest(double v)
{
    uint64_t t = (uint64_t)v;
    return (v - (double)t > 0.5) ? t + 1 : t;
}

// ---------------------------------------------------------------------------
// k_margin: one wave per sample row
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_margin(const double* __restrict__ A, const double* __restrict__ x,
                                                int m, int m_pad, int d, int d_pad, double inv_m,
                                                double* __restrict__ Dw, double* __restrict__ gw,
                                                double* __restrict__ nlog)
{
    extern __shared__ double xs[];
    const int c = blockIdx.y;
    for (int i = threadIdx.x; i < d_pad; i += blockDim.x)
        xs[i] = (i < d) ? x[i] : 0.0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s0 = blockIdx.x * 16 + wave * 4;
    for (int q = 0; q < 4; ++q) {
        const int s = s0 + q;
        if (s >= m_pad)
            break;
        const double* row = A + ((size_t)c * m_pad + s) * d_pad;
        double acc = 0.0;
        for (int i = lane; i < d_pad; i += 64)
            acc = fma(row[i], xs[i], acc);
        for (int off = 32; off > 0; off >>= 1)
            acc += __shfl_xor(acc, off);
        if (lane == 0) {
            const size_t o = (size_t)c * m_pad + s;
            if (s < m) {
                // logistic_regression.h:98-101,148-161,398-404
                const double sg = 1.0 / (1.0 + exp(-acc));
                const double z1 = 1.0 - sg;
                const double z2 = z1 * sg;
                Dw[o] = inv_m * z2;
                gw[o] = z1 * (-inv_m);
                nlog[o] = -log(sg);
            } else {
                Dw[o] = 0.0;
                gw[o] = 0.0;
                nlog[o] = 0.0;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_hessian_f64: H_c = A_c^T diag(D) A_c on the upper tiles, FP64 MFMA
// block = 4 waves (2 x 2), each wave a 32 x 32 sub-tile of 2 x 2 MFMA 16x16 blocks.
// Blocks are mapped so that all tiles of one client run on one XCD
// (blockIdx % 8 selects the XCD), keeping the client's rows of A in that L2.
// ---------------------------------------------------------------------------
template <bool kInit>
__global__ __launch_bounds__(256) void k_hessian_f64(const double* __restrict__ A, const double* __restrict__ Dw,
                                                     const double* __restrict__ gw, const double* __restrict__ x,
                                                     int n, int m_pad, int d, int d_pad, int tiles_up,
                                                     double lambda, size_t S, double* __restrict__ learned,
                                                     double* __restrict__ delta, double* __restrict__ grad_c,
                                                     double* __restrict__ part)
{
    const int b = blockIdx.x;
    const int xcd = b & 7, slot = b >> 3;
    const int q = slot / tiles_up, t = slot - q * tiles_up;
    const int c = xcd + 8 * q;
    if (c >= n)
        return;
    int bj = 0;
    while ((bj + 1) * (bj + 2) / 2 <= t) ++bj;
    const int bi = t - bj * (bj + 1) / 2;
    const int i0 = bi * kTile, j0 = bj * kTile;
    const bool dg = (bi == bj);

    __shared__ double sI[kKC][kTile + kPadLds];
    __shared__ double sJ[kKC][kTile + kPadLds];
    __shared__ double sD[kKC];
    __shared__ double sG[kKC];
    __shared__ double red[2][4];
    __shared__ double gred[4][kTile];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave & 1, wj = wave >> 1;
    const int l16 = lane & 15, kq = lane >> 4;

    dbl4 acc[2][2];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int it = 0; it < 2; ++it) acc[jt][it] = dbl4{0.0, 0.0, 0.0, 0.0};
    double gacc = 0.0;

    const double* Ac = A + (size_t)c * m_pad * d_pad;
    const double* Dc = Dw + (size_t)c * m_pad;
    const double* Gc = gw + (size_t)c * m_pad;

    for (int s0 = 0; s0 < m_pad; s0 += kKC) {
        if (tid < kKC) {
            sD[tid] = Dc[s0 + tid];
            sG[tid] = Gc[s0 + tid];
        }
        __syncthreads();
        for (int e = tid; e < kKC * kTile; e += 256) {
            const int kk = e >> 6, ii = e & 63;
            const double* row = Ac + (size_t)(s0 + kk) * d_pad;
            const double a = row[i0 + ii];
            sI[kk][ii] = a;
            const double bv = dg ? a : row[j0 + ii];
            sJ[kk][ii] = sD[kk] * bv;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kKC; ks += 4) {
            const int kk = ks + kq;
            const double a0 = sJ[kk][wj * 32 + l16];
            const double a1 = sJ[kk][wj * 32 + 16 + l16];
            const double b0 = sI[kk][wi * 32 + l16];
            const double b1 = sI[kk][wi * 32 + 16 + l16];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (!kInit && dg) {
            const int ii = tid & 63, g4 = tid >> 6;
#pragma unroll
            for (int kk = g4 * 4; kk < g4 * 4 + 4; ++kk)
                gacc = fma(sI[kk][ii], sG[kk], gacc);
        }
        __syncthreads();
    }

    // epilogue: lane (l16, kq), item r of v_mfma_f64_16x16x4f64 holds
    // D[m = 4r + kq][n = l16] (pinned by the layout probe test), i.e.
    // rows i = .. + l16 and columns j = .. + kq + 4r of the sub-tile (jt, it)
    double soff = 0.0, sdiag = 0.0;
#pragma unroll
    for (int jt = 0; jt < 2; ++jt) {
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const dbl4 v = acc[jt][it];
            const int i = i0 + wi * 32 + it * 16 + l16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + wj * 32 + jt * 16 + kq + 4 * r;
                if (j < d && i <= j) {
                    double h = v[r];
                    if (i == j)
                        h = h + lambda;
                    const size_t o = (size_t)c * S + (size_t)j * (j + 1) / 2 + i;
                    if (kInit) {
                        learned[o] = h;
                    } else {
                        const double dv = h - learned[o];
                        delta[o] = dv;
                        if (i == j)
                            sdiag = fma(dv, dv, sdiag);
                        else
                            soff = fma(dv, dv, soff);
                    }
                }
            }
        }
    }
    if (kInit)
        return;
    for (int off = 32; off > 0; off >>= 1) {
        soff += __shfl_xor(soff, off);
        sdiag += __shfl_xor(sdiag, off);
    }
    if (lane == 0) {
        red[0][wave] = soff;
        red[1][wave] = sdiag;
    }
    if (dg)
        gred[tid >> 6][tid & 63] = gacc;
    __syncthreads();
    if (tid == 0) {
        const double o = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        const double dd = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        part[((size_t)c * tiles_up + t) * 2 + 0] = o;
        part[((size_t)c * tiles_up + t) * 2 + 1] = dd;
    }
    if (dg && tid < kTile) {
        const int i = i0 + tid;
        if (i < d) {
            const double s = (gred[0][tid] + gred[1][tid]) + (gred[2][tid] + gred[3][tid]);
            grad_c[(size_t)c * d_pad + i] = lambda * x[i] + s;
        }
    }
}

// ---------------------------------------------------------------------------
// dopt's MinHeap<4>::minHeapify (MinHeap.h:255-290), values and ids in LDS.
// After a swap it keeps comparing against the value that moved up.
// ---------------------------------------------------------------------------
__device__ inline void heapify(double* hv, uint32_t* hid, int i, int n)
{
    double lowv = hv[i];
    for (;;) {
        int lowest = i;
        const int first = 4 * i + 1;
        const int bound = min(first + 4, n);
        for (int ch = first; ch < bound; ++ch) {
            const double cv = hv[ch];
            if (lowv > cv) {
                lowest = ch;
                lowv = cv;
            }
        }
        if (lowest != i) {
            const double tv = hv[i];
            hv[i] = hv[lowest];
            hv[lowest] = tv;
            const uint32_t ti = hid[i];
            hid[i] = hid[lowest];
            hid[lowest] = ti;
            i = lowest;
        } else {
            break;
        }
    }
}

__device__ inline int heap_level_first(int L)
{
    int f = 0, w = 1;
    for (int q = 0; q < L; ++q) {
        f += w;
        w *= 4;
    }
    return f;
}

// in-place ascending bitonic sort of P (power of two) u32 keys in LDS
__device__ inline void bitonic_sort_u32(uint32_t* keys, int P)
{
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < (P >> 1); i += blockDim.x) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const uint32_t a = keys[lo], bb = keys[hi];
                if ((a > bb) == up) {
                    keys[lo] = bb;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

__device__ inline void bitonic_sort_u64(uint64_t* keys, int P)
{
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < (P >> 1); i += blockDim.x) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const uint64_t a = keys[lo], bb = keys[hi];
                if ((a > bb) == up) {
                    keys[lo] = bb;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

__host__ __device__ inline int pow2_at_least(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Lk (TopK/TopLEK worker: frobeniusNormForSymmetricMatrixFromUpPart of Delta,
// LocalWorkerForFedNL_TopK.h:74-78) from the Hessian kernel's tile partials,
// summed in tile order.
__device__ inline double fro2_from_parts(const double* part, int c, int tiles_up)
{
    double o = 0.0, dd = 0.0;
    for (int t = 0; t < tiles_up; ++t) {
        o = o + part[((size_t)c * tiles_up + t) * 2 + 0];
        dd = dd + part[((size_t)c * tiles_up + t) * 2 + 1];
    }
    return (o + o) + dd;
}

// ---------------------------------------------------------------------------
// k_select_heap: dopt's getTopKFromUpperDiagonalPart<true> heap scan
// (SpecialMathRoutinesForMatrix.h:276-327) and, for TopLEK, the energy walk of
// getTopLEKFromUpperDiagonalPart<true> (:377-511) up to the random draw.
// One workgroup per client. Wave 0 runs the sequential scan: each chunk of 64
// list positions is tested against the current root with one ballot; only
// admitted positions enter the (sequential) heap update. Waves 1-3 stage the
// next segment of |Delta| into LDS meanwhile.
// Output: sorted selected list positions, count; TopLEK: hit flag, p, popped.
// ---------------------------------------------------------------------------
template <bool kLek>
__global__ __launch_bounds__(256) void k_select_heap(const double* __restrict__ delta, size_t S, int k,
                                                     const double* __restrict__ part, int tiles_up,
                                                     double lek_delta, double* __restrict__ lk_out,
                                                     uint32_t* __restrict__ sel, int* __restrict__ sel_cnt,
                                                     double* __restrict__ lek_p, int* __restrict__ lek_hit,
                                                     uint32_t* __restrict__ lek_popped)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = blockIdx.x;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double* dc = delta + (size_t)c * S;
    __shared__ double s_fro2;
    if (tid == 0) {
        const double f2 = part ? fro2_from_parts(part, c, tiles_up) : 0.0;
        s_fro2 = f2;
        if (lk_out)
            lk_out[c] = sqrt(f2);
    }
    if ((size_t)k >= S) {
        for (size_t p = tid; p < S; p += blockDim.x)
            sel[(size_t)c * k + p] = (uint32_t)p;
        if (tid == 0) {
            sel_cnt[c] = (int)S;
            if (kLek)
                lek_hit[c] = 0;
        }
        return;
    }
    const int P = pow2_at_least(k);
    double* hv = (double*)smem;
    uint32_t* hid = (uint32_t*)(hv + k);
    double* seg = (double*)(hid + round_up(P, 2));

    for (int q = tid; q < k; q += blockDim.x) {
        hv[q] = fabs(dc[q]);
        hid[q] = (uint32_t)q;
    }
    __syncthreads();

    // buildMinHeap (MinHeap.h:147-191): nodes (k+2)/4 .. 0, one level at a time
    // (same-level nodes have disjoint subtrees).
    {
        const int i0 = (k + 2) / 4;
        int L = 0;
        while (heap_level_first(L + 1) <= i0) ++L;
        for (int lvl = L; lvl >= 0; --lvl) {
            const int lo = heap_level_first(lvl);
            const int hi = min(heap_level_first(lvl + 1) - 1, i0);
            for (int nd = lo + tid; nd <= hi; nd += blockDim.x)
                heapify(hv, hid, nd, k);
            __syncthreads();
        }
    }

    // scan of list positions k .. S-1
    const size_t rest = S - (size_t)k;
    const int nseg = (int)((rest + kSeg - 1) / kSeg);
    for (int e = tid; e < kSeg && (size_t)e < rest; e += blockDim.x)
        seg[e] = fabs(dc[(size_t)k + e]);
    __syncthreads();
    double rv = hv[0];
    double rv0 = rv;
    for (int sgi = 0; sgi < nseg; ++sgi) {
        double* cur = seg + (sgi & 1) * kSeg;
        const size_t sbase = (size_t)k + (size_t)sgi * kSeg;
        const int slen = (int)min((size_t)kSeg, S - sbase);
        if (wave == 0) {
            for (int e0 = 0; e0 < slen; e0 += 64) {
                const int e = e0 + lane;
                const double v = (e < slen) ? cur[e] : -1.0;
                uint64_t mask = __ballot(v > rv);
                while (mask) {
                    const int f = __builtin_ctzll(mask);
                    const double key = __shfl(v, f);
                    if (lane == 0) {
                        hv[0] = key;  // increaseMinimumInMinHeap (MinHeap.h:78-85)
                        hid[0] = (uint32_t)(sbase + e0 + f);
                        heapify(hv, hid, 0, k);
                        rv0 = hv[0];
                    }
                    rv = __shfl(rv0, 0);
                    const uint64_t above = (f == 63) ? 0ull : (~0ull << (f + 1));
                    mask = __ballot(v > rv) & above;
                }
            }
        } else if (sgi + 1 < nseg) {
            double* nxt = seg + ((sgi + 1) & 1) * kSeg;
            const size_t nbase = sbase + kSeg;
            const int nlen = (int)min((size_t)kSeg, S - nbase);
            for (int e = tid - 64; e < nlen; e += blockDim.x - 64)
                nxt[e] = fabs(dc[nbase + e]);
        }
        __syncthreads();
    }

    int cnt = k;
    if (kLek) {
        __shared__ int s_cnt;
        if (tid == 0) {
            // SpecialMathRoutinesForMatrix.h:383-511
            const double fro2 = s_fro2;
            const double one_minus_delta_xf = (1.0 - lek_delta) * fro2;
            const double eps = 1e-1 * fro2;
            double numerator = fro2;
            for (int q = 0; q < k; ++q) {
                const double it = hv[q];
                if (pos_is_diag(hid[q]))
                    numerator -= it * it;
                else
                    numerator -= 2.0 * it * it;
            }
            int nh = k;
            int hit = 0;
            double pp = 0.0;
            uint32_t popped = 0;
            do {
                const double prev = numerator;
                const uint32_t pid = hid[0];
                const double pv = hv[0];
                // extracFromMinHeap (MinHeap.h:125-141)
                hv[0] = hv[nh - 1];
                hid[0] = hid[nh - 1];
                hv[nh - 1] = pv;
                hid[nh - 1] = pid;
                heapify(hv, hid, 0, nh - 1);
                nh -= 1;
                if (pos_is_diag(pid))
                    numerator += pv * pv;
                else
                    numerator += 2.0 * pv * pv;
                const double next = numerator;
                if (one_minus_delta_xf + eps > prev && one_minus_delta_xf < next + eps) {
                    pp = (one_minus_delta_xf - prev + eps) / (next - prev + eps);
                    hit = 1;
                    popped = pid;
                    break;
                }
            } while (nh != 0);
            s_cnt = nh;
            lek_hit[c] = hit;
            lek_p[c] = pp;
            lek_popped[c] = popped;
        }
        __syncthreads();
        cnt = s_cnt;
    }

    for (int q = cnt + tid; q < P; q += blockDim.x)
        hid[q] = 0xFFFFFFFFu;
    __syncthreads();
    bitonic_sort_u32(hid, P);
    for (int q = tid; q < cnt; q += blockDim.x)
        sel[(size_t)c * k + q] = hid[q];
    if (tid == 0)
        sel_cnt[c] = cnt;
}

// ---------------------------------------------------------------------------
// k_select_randk: generateRandKItemsInUpperTriangularPart
// (SpecialMathRoutinesForMatrix.h:95-107, Shuffle.h:31-93) without the
// sequential swaps: every draw u_i comes from an LCG jump-ahead, the swap
// target t_i = i + round(u_i (S-1-i)); the value left at slot i is found by
// following the chain of earlier swaps that touched its source slot.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_select_randk(size_t S, int k, uint32_t* __restrict__ rng,
                                                      uint32_t* __restrict__ sel, int* __restrict__ sel_cnt)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const int P = pow2_at_least(k);
    uint64_t* keys = (uint64_t*)smem;            // P
    int* qprev = (int*)(keys + P);                // k
    int* pprev = qprev + k;                       // k
    uint32_t* outk = (uint32_t*)keys;             // reused after the chains are resolved
    const uint32_t x0 = rng[c];

    for (int i = tid; i < P; i += blockDim.x) {
        if (i < k) {
            uint32_t Aj, Cj;
            lcg_jump((uint64_t)i + 1, Aj, Cj);
            const uint32_t xi = Aj * x0 + Cj;
            const double u = (double)xi * (1.0 / 4294967295.0);
            const uint64_t off = round_nearest(u * (double)(S - 1 - (size_t)i));
            const uint64_t ti = (uint64_t)i + off;
            keys[i] = (ti << 32) | (uint64_t)i;
            qprev[i] = -1;
            pprev[i] = -1;
        } else {
            keys[i] = ~0ull;
        }
    }
    __syncthreads();
    bitonic_sort_u64(keys, P);
    // qprev[i]: last j < i with t_j == t_i; pprev[j]: last j' < j with t_j' == j
    for (int q = tid; q < k; q += blockDim.x) {
        const uint64_t key = keys[q];
        const int i = (int)(key & 0xFFFFFFFFull);
        const uint64_t ti = key >> 32;
        if (q > 0 && (keys[q - 1] >> 32) == ti)
            qprev[i] = (int)(keys[q - 1] & 0xFFFFFFFFull);
        if (ti < (uint64_t)k) {
            // every j with t_j == ti and j < ti: the largest such precedes (ti, ti) in order
            // this entry is one of them when it is the last with index < ti
            const int j = (int)ti;
            if ((uint64_t)i < ti) {
                const bool last = (q + 1 == k) || ((keys[q + 1] >> 32) != ti) ||
                                  ((keys[q + 1] & 0xFFFFFFFFull) >= ti);
                if (last)
                    pprev[j] = i;
            }
        }
    }
    __syncthreads();
    // resolve: out[i] = value at slot i after step i (keys are dead from here on)
    for (int i = tid; i < k; i += blockDim.x) {
        uint32_t Aj, Cj;
        lcg_jump((uint64_t)i + 1, Aj, Cj);
        const uint32_t xi = Aj * x0 + Cj;
        const double u = (double)xi * (1.0 / 4294967295.0);
        const uint64_t ti = (uint64_t)i + round_nearest(u * (double)(S - 1 - (size_t)i));
        int src;
        if (ti == (uint64_t)i)
            src = i;
        else
            src = qprev[i];
        uint64_t v;
        if (src < 0) {
            v = ti;
        } else {
            int r = src;
            while (pprev[r] >= 0) r = pprev[r];
            v = (uint64_t)r;
        }
        outk[i] = (uint32_t)v;
    }
    for (int i = k + tid; i < P; i += blockDim.x)
        outk[i] = 0xFFFFFFFFu;
    __syncthreads();
    bitonic_sort_u32(outk, P);
    for (int q = tid; q < k; q += blockDim.x)
        sel[(size_t)c * k + q] = outk[q];
    if (tid == 0) {
        uint32_t Ak, Ck;
        lcg_jump((uint64_t)k, Ak, Ck);
        rng[c] = Ak * x0 + Ck;
        sel_cnt[c] = k;
    }
}

// RandSeqK (SpecialMathRoutinesForMatrix.h:137-169): window of k list positions
// from generateInteger() % S, cyclic; stored ascending. Identical: all positions.
__global__ __launch_bounds__(256) void k_select_window(int comp, size_t S, int k, uint32_t* __restrict__ rng,
                                                       uint32_t* __restrict__ sel, int* __restrict__ sel_cnt,
                                                       uint32_t* __restrict__ seq_start)
{
    const int c = blockIdx.x;
    __shared__ uint64_t s_start;
    if (threadIdx.x == 0) {
        uint64_t st = 0;
        if (comp == FEDNL_RANDSEQK) {
            const uint32_t xv = kLcgA * rng[c] + kLcgC;
            rng[c] = xv;
            st = (uint64_t)xv % S;
        }
        s_start = st;
        if (seq_start)
            seq_start[c] = (uint32_t)st;
        sel_cnt[c] = (comp == FEDNL_RANDSEQK) ? k : (int)S;
    }
    __syncthreads();
    if (comp == FEDNL_RANDSEQK) {
        const uint64_t st = s_start;
        const uint64_t wrap = (st + (uint64_t)k > S) ? st + (uint64_t)k - S : 0;  // leading run [0, wrap)
        for (int t = threadIdx.x; t < k; t += blockDim.x) {
            const uint64_t p = ((uint64_t)t < wrap) ? (uint64_t)t : st + ((uint64_t)t - wrap);
            sel[(size_t)c * k + t] = (uint32_t)p;
        }
    } else {
        for (size_t p = threadIdx.x; p < S; p += blockDim.x)
            sel[(size_t)c * k + p] = (uint32_t)p;
    }
}

// ---------------------------------------------------------------------------
// k_emit: message values and learned-Hessian update of every client
// (LocalWorkerForFedNL_TopK.h:89-129; RandK/RandSeqK :127-162 with the
// S/k multiplier; TopLEK :93-110). TopLEK clients that hit the rounding window
// draw from their worker's RealLinear stream: the draw index is the number of
// earlier hits among the clients of the same worker this round.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_emit(int comp, const double* __restrict__ delta, size_t S, int k, int lda,
                                              const uint32_t* __restrict__ sel, const int* __restrict__ sel_cnt,
                                              const double* __restrict__ lek_p, const int* __restrict__ lek_hit,
                                              const uint32_t* __restrict__ lek_popped,
                                              const uint32_t* __restrict__ lek_state, int workers, double alpha,
                                              double prescale, double mult, double* __restrict__ learned,
                                              uint32_t* __restrict__ pos_out, uint32_t* __restrict__ idx_out,
                                              double* __restrict__ vals_out, int* __restrict__ cnt_out)
{
    const int c = blockIdx.x;
    __shared__ int s_cnt, s_ins;
    __shared__ uint32_t s_pop;
    if (threadIdx.x == 0) {
        int cnt = sel_cnt[c];
        int ins = -1;
        uint32_t popped = 0;
        if (comp == FEDNL_TOPLEK && lek_hit[c]) {
            const int w = c % workers;
            int h = 0;
            for (int cc = w; cc < c; cc += workers) h += lek_hit[cc];
            uint32_t st = lek_state[w];
            for (int q = 0; q <= h; ++q) st = lcg_real_step(st);
            const double u = (double)st * (1.0 / 4294967295.0);
            if (u > lek_p[c]) {
                // insertInMinHeap then sort: the popped position joins the set
                popped = lek_popped[c];
                int lo = 0, hi = cnt;
                const uint32_t* sc = sel + (size_t)c * k;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (sc[mid] < popped) lo = mid + 1; else hi = mid;
                }
                ins = lo;
                cnt += 1;
            }
        }
        s_cnt = cnt;
        s_ins = ins;
        s_pop = popped;
        cnt_out[c] = cnt;
    }
    __syncthreads();
    const int cnt = s_cnt, ins = s_ins;
    const bool scaled = (comp == FEDNL_RANDK || comp == FEDNL_RANDSEQK);
    const uint32_t* sc = sel + (size_t)c * k;
    const double* dc = delta + (size_t)c * S;
    double* lc = learned + (size_t)c * S;
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
        uint32_t p;
        if (ins < 0 || t < ins) p = sc[t];
        else if (t == ins) p = s_pop;
        else p = sc[t - 1];
        const double dv = dc[p];
        const double item = scaled ? dv * mult : dv;
        vals_out[(size_t)c * k + t] = item * prescale;
        const double upd = alpha * item;
        lc[p] = lc[p] + upd;
        pos_out[(size_t)c * k + t] = p;
        idx_out[(size_t)c * k + t] = pos_flat(p, lda);
    }
    for (int t = cnt + threadIdx.x; t < k; t += blockDim.x) {
        pos_out[(size_t)c * k + t] = 0xFFFFFFFFu;
        idx_out[(size_t)c * k + t] = 0xFFFFFFFFu;
        vals_out[(size_t)c * k + t] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// k_master_h: H[idx] += val for clients 0..n-1 in order (TrainLoop.h:1197-1274).
// One workgroup per window of list positions; client ranges by binary search.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_master_h(double* __restrict__ H, size_t S, int n, int k,
                                                  const uint32_t* __restrict__ pos, const double* __restrict__ vals,
                                                  const int* __restrict__ cnt)
{
    __shared__ double hw[kWin];
    __shared__ int lo_s[kMaxClients], hi_s[kMaxClients];
    const size_t p0 = (size_t)blockIdx.x * kWin;
    const size_t p1 = min(p0 + (size_t)kWin, S);
    for (size_t i = threadIdx.x; i < p1 - p0; i += blockDim.x)
        hw[i] = H[p0 + i];
    for (int cc = threadIdx.x; cc < n; cc += blockDim.x) {
        const uint32_t* pc = pos + (size_t)cc * k;
        const int m = cnt[cc];
        int a = 0, b = m;
        while (a < b) { const int mid = (a + b) >> 1; if (pc[mid] < p0) a = mid + 1; else b = mid; }
        int a2 = a, b2 = m;
        while (a2 < b2) { const int mid = (a2 + b2) >> 1; if (pc[mid] < p1) a2 = mid + 1; else b2 = mid; }
        lo_s[cc] = a;
        hi_s[cc] = a2;
    }
    __syncthreads();
    for (int cc = 0; cc < n; ++cc) {
        const int a = lo_s[cc], b = hi_s[cc];
        if (a == b)
            continue;
        const uint32_t* pc = pos + (size_t)cc * k;
        const double* vc = vals + (size_t)cc * k;
        for (int e = a + threadIdx.x; e < b; e += blockDim.x) {
            const size_t o = pc[e] - p0;
            hw[o] = hw[o] + vc[e];
        }
        __syncthreads();
    }
    for (size_t i = threadIdx.x; i < p1 - p0; i += blockDim.x)
        H[p0 + i] = hw[i];
}

// H = (sum_c learned_c) / n, clients in order (TrainLoop.h:639-643, 732-735)
__global__ __launch_bounds__(256) void k_init_H(double* __restrict__ H, const double* __restrict__ learned, size_t S,
                                                int n, double inv_n)
{
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S)
        return;
    double s = 0.0;
    for (int c = 0; c < n; ++c)
        s = s + learned[(size_t)c * S + p];
    H[p] = s * inv_n;
}

__device__ inline double block_sum(double v, double* scratch)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0)
        scratch[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s = s + scratch[w];
    return s;
}

// ---------------------------------------------------------------------------
// k_master_vec: averaged gradient, Lk, f (TrainLoop.h:1100-1124,1278-1333,
// 1512-1544), norms for the record (TrainLoop.h:1679-1691), TopLEK stream
// advance. One workgroup.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_master_vec(int comp, int n, int d, int d_pad, int m, int m_pad,
                                                    const double* __restrict__ grad_c, const double* __restrict__ part,
                                                    int tiles_up, double* __restrict__ lk,
                                                    const double* __restrict__ nlog, const double* __restrict__ x,
                                                    double lambda, double inv_n, int option_b, int tracking,
                                                    const int* __restrict__ lek_hit, uint32_t* __restrict__ lek_state,
                                                    int workers, double* __restrict__ g_avg,
                                                    double* __restrict__ shift_out, double* __restrict__ trace,
                                                    const int* __restrict__ round_ctr, int trace_cap)
{
    __shared__ double scratch[16];
    double gg = 0.0, xx = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < n; ++c)
            s = s + grad_c[(size_t)c * d_pad + i];
        const double g = s * inv_n;
        g_avg[i] = g;
        gg = fma(g, g, gg);
        xx = fma(x[i], x[i], xx);
    }
    gg = block_sum(gg, scratch);
    xx = block_sum(xx, scratch);
    for (int c = threadIdx.x; c < n; c += blockDim.x)
        lk[c] = sqrt(fro2_from_parts(part, c, tiles_up));
    __syncthreads();
    double fsum = 0.0;
    if (tracking) {
        // f_c = mean_s(-log sigma_s) + lambda/2 ||x||^2 (logistic_regression.h:119-123)
        __shared__ double fc_s[kMaxClients];
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            double s = 0.0;
            for (int q = 0; q < m; ++q)
                s = s + nlog[(size_t)c * m_pad + q];
            fc_s[c] = s / (double)m + (lambda / 2.0) * xx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int c = 0; c < n; ++c) fsum = fsum + fc_s[c];
            fsum = fsum * inv_n;
        }
    }
    if (threadIdx.x == 0) {
        double l = 0.0;
        if (lk)
            for (int c = 0; c < n; ++c) l = l + lk[c];
        l = l * inv_n;
        *shift_out = option_b ? l : lambda;  // option a: mu(f) = lambda (logistic_regression.h:702-705)
        const int r = *round_ctr;
        if (r < trace_cap) {
            trace[(size_t)r * 4 + 0] = sqrt(gg);
            trace[(size_t)r * 4 + 1] = fsum;
            trace[(size_t)r * 4 + 2] = sqrt(xx);
        }
    }
    if (comp == FEDNL_TOPLEK) {
        for (int w = threadIdx.x; w < workers; w += blockDim.x) {
            uint32_t st = lek_state[w];
            for (int c = w; c < n; c += workers)
                if (lek_hit[c]) st = lcg_real_step(st);
            lek_state[w] = st;
        }
    }
}

// ---------------------------------------------------------------------------
// k_newton: Cholesky of (H + shift I) and the two triangular solves of the
// master (TrainLoop.h:1449-1466; CholeskyFactorization.h:153-272;
// ElementarySolvers.h:72-226). Right-looking blocked factorisation in one
// workgroup: each 32-column panel is factored in LDS, the trailing lower
// triangle is updated in L2-resident scratch W (column-major lower, ld = dp).
// On success x -= dx. The factorisation fails, as dopt's, when a pivot is < 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_newton(int d, int dp, const double* __restrict__ Hup,
                                                 const double* __restrict__ shift_p, const double* __restrict__ b,
                                                 double* __restrict__ W, double* __restrict__ x,
                                                 double* __restrict__ dx_out, double* __restrict__ trace,
                                                 int* __restrict__ round_ctr, int trace_cap, int update_x)
{
    extern __shared__ __align__(16) double lds[];
    double* Pn = lds;                // kNB x dp (column cc at Pn + cc*dp, rows relative to jb)
    __shared__ double yv[kMaxNewtonD];
    __shared__ int s_fail;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const double shift = *shift_p;
    if (tid == 0)
        s_fail = 0;
    // W(i,j) = H(j,i) (+shift on the diagonal) for i >= j
    for (size_t e = tid; e < (size_t)d * d; e += nthr) {
        const int j = (int)(e / d), i = (int)(e % d);
        if (i >= j) {
            double v = Hup[(size_t)i * (i + 1) / 2 + j];
            if (i == j)
                v = v + shift;
            W[(size_t)j * dp + i] = v;
        }
    }
    __syncthreads();

    for (int jb = 0; jb < d; jb += kNB) {
        const int nb = min(kNB, d - jb);
        const int R = d - jb;
        for (int e = tid; e < nb * R; e += nthr) {
            const int cc = e / R, r = e % R;
            if (r >= cc)
                Pn[cc * dp + r] = W[(size_t)(jb + cc) * dp + jb + r];
        }
        __syncthreads();
        for (int cc = 0; cc < nb; ++cc) {
            if (tid == 0) {
                const double v = Pn[cc * dp + cc];
                if (v < 0.0)
                    s_fail = 1;
                Pn[cc * dp + cc] = sqrt(v);
            }
            __syncthreads();
            if (s_fail)
                break;
            const double inv = 1.0 / Pn[cc * dp + cc];
            for (int r = cc + 1 + tid; r < R; r += nthr)
                Pn[cc * dp + r] = inv * Pn[cc * dp + r];
            __syncthreads();
            // rank-1 update of the remaining panel columns
            const int w = nb - cc - 1;
            if (w > 0) {
                for (int e = tid; e < w * R; e += nthr) {
                    const int c2 = cc + 1 + e / R, r = e % R;
                    if (r >= c2)
                        Pn[c2 * dp + r] = fma(-Pn[cc * dp + r], Pn[cc * dp + c2], Pn[c2 * dp + r]);
                }
            }
            __syncthreads();
        }
        if (s_fail)
            break;
        for (int e = tid; e < nb * R; e += nthr) {
            const int cc = e / R, r = e % R;
            if (r >= cc)
                W[(size_t)(jb + cc) * dp + jb + r] = Pn[cc * dp + r];
        }
        // trailing update of the lower triangle, 4 x 4 register tiles
        const int T = R - nb;
        if (T > 0) {
            const int tb = (T + 3) / 4;
            const int ntiles = tb * (tb + 1) / 2;
            for (int tt = tid; tt < ntiles; tt += nthr) {
                int ti = (int)((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
                while (ti * (ti + 1) / 2 > tt) --ti;
                while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
                const int tj = tt - ti * (ti + 1) / 2;  // tj <= ti
                const int rb = nb + ti * 4, cb = nb + tj * 4;  // rows i, cols j (relative to jb)
                double a4[4][4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) a4[u][v] = 0.0;
                for (int cc = 0; cc < nb; ++cc) {
                    double ri[4], cj[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        ri[u] = (rb + u < R) ? Pn[cc * dp + rb + u] : 0.0;
                        cj[u] = (cb + u < R) ? Pn[cc * dp + cb + u] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) a4[u][v] = fma(ri[u], cj[v], a4[u][v]);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = cb + v;
                    if (j >= R)
                        continue;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int i = rb + u;
                        if (i < R && i >= j) {
                            double* wp = &W[(size_t)(jb + j) * dp + jb + i];
                            *wp = *wp - a4[u][v];
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

    const int r = *round_ctr;
    if (s_fail) {
        if (tid == 0) {
            if (r < trace_cap)
                trace[(size_t)r * 4 + 3] = 0.0;
            *round_ctr = r + 1;
        }
        for (int i = tid; i < d; i += nthr)
            dx_out[i] = 0.0;
        return;
    }

    // forward: L y = b
    for (int i = tid; i < d; i += nthr)
        yv[i] = b[i];
    __syncthreads();
    for (int jb = 0; jb < d; jb += kNB) {
        const int nb = min(kNB, d - jb);
        if (tid < 64) {
            for (int q = 0; q < nb; ++q) {
                const int i = jb + q;
                if (tid == 0)
                    yv[i] = yv[i] / W[(size_t)i * dp + i];
                __builtin_amdgcn_wave_barrier();
                const double yi = yv[i];
                const int rr = i + 1 + tid;
                if (rr < jb + nb)
                    yv[rr] = fma(-W[(size_t)i * dp + rr], yi, yv[rr]);
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        for (int rr = jb + nb + tid; rr < d; rr += nthr) {
            double s = yv[rr];
            for (int q = 0; q < nb; ++q)
                s = fma(-W[(size_t)(jb + q) * dp + rr], yv[jb + q], s);
            yv[rr] = s;
        }
        __syncthreads();
    }
    // backward: L^T x = y
    const int last_jb = ((d - 1) / kNB) * kNB;
    for (int jb = last_jb; jb >= 0; jb -= kNB) {
        const int nb = min(kNB, d - jb);
        if (tid < 64) {
            for (int q = nb - 1; q >= 0; --q) {
                const int i = jb + q;
                if (tid == 0)
                    yv[i] = yv[i] / W[(size_t)i * dp + i];
                __builtin_amdgcn_wave_barrier();
                const double xi = yv[i];
                const int rr = jb + tid;
                if (rr < i)
                    yv[rr] = fma(-W[(size_t)rr * dp + i], xi, yv[rr]);
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        for (int rr = tid; rr < jb; rr += nthr) {
            double s = yv[rr];
            for (int q = 0; q < nb; ++q)
                s = fma(-W[(size_t)rr * dp + jb + q], yv[jb + q], s);
            yv[rr] = s;
        }
        __syncthreads();
    }
    for (int i = tid; i < d; i += nthr) {
        dx_out[i] = yv[i];
        if (update_x)
            x[i] = x[i] - yv[i];
    }
    if (tid == 0) {
        if (r < trace_cap)
            trace[(size_t)r * 4 + 3] = 1.0;
        *round_ctr = r + 1;
    }
}

// Lane layout check of v_mfma_f64_16x16x4f64: with A supplied as (lane+1) and
// B as 1 (and the reverse), each output element reveals which lanes fed its
// row / column. One wave; out holds 3 x 64 x 4 doubles.
__global__ __launch_bounds__(64) void k_mfma_probe(double* __restrict__ out)
{
    const int lane = threadIdx.x;
    const dbl4 z = {0.0, 0.0, 0.0, 0.0};
    const dbl4 r1 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(lane + 1), 1.0, z, 0, 0, 0);
    const dbl4 r2 = __builtin_amdgcn_mfma_f64_16x16x4f64(1.0, (double)(lane + 1), z, 0, 0, 0);
    const dbl4 r3 = __builtin_amdgcn_mfma_f64_16x16x4f64((lane < 16) ? (double)(lane + 1) : 0.0,
                                                         (lane < 16) ? 1.0 : 0.0, z, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        out[lane * 4 + q] = r1[q];
        out[256 + lane * 4 + q] = r2[q];
        out[512 + lane * 4 + q] = r3[q];
    }
}

// ---------------------------------------------------------------------------
// host engine
// ---------------------------------------------------------------------------
enum KernelId { KID_MARGIN, KID_HESS, KID_SELECT, KID_EMIT, KID_MASTER_H, KID_MASTER_VEC, KID_NEWTON, KID_COUNT };
static const char* kKernelNames[KID_COUNT] = {"k_margin", "k_hessian_f64", "k_select", "k_emit",
                                              "k_master_h", "k_master_vec", "k_newton"};

struct Engine {
    fednl_opts o;
    int n = 0, m = 0, d = 0, k = 0, lda = 0, m_pad = 0, d_pad = 0, nt = 0, tiles_up = 0, dp = 0, workers = 1;
    size_t S = 0;
    double alpha = 1.0, prescale = 0.0, inv_n = 0.0, mult = 1.0, lek_delta = 0.0;
    hipStream_t st = nullptr;

    double *A = nullptr, *x = nullptr, *Dw = nullptr, *gw = nullptr, *nlog = nullptr;
    double *learned = nullptr, *delta = nullptr, *grad_c = nullptr, *part = nullptr, *lk = nullptr;
    uint32_t *sel = nullptr, *pos = nullptr, *idx = nullptr, *rng = nullptr, *lek_state = nullptr, *lek_popped = nullptr;
    uint32_t *seq_start = nullptr;
    int *sel_cnt = nullptr, *cnt = nullptr, *lek_hit = nullptr, *round_ctr = nullptr;
    double *lek_p = nullptr, *vals = nullptr, *H = nullptr, *g_avg = nullptr, *shift = nullptr, *W = nullptr;
    double *dx = nullptr, *trace = nullptr;
    int trace_cap = 0;
    int rounds_enqueued = 0;

    hipEvent_t ev_start = nullptr;
    std::vector<hipEvent_t> round_ev;
    bool timing = false;
    struct Pending { int kid; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;
    double kms[KID_COUNT] = {0};
    int klaunch[KID_COUNT] = {0};
    size_t lds_select = 0, lds_newton = 0;
};

static hipEvent_t pool_event(Engine* e)
{
    if (!e->ev_pool.empty()) {
        hipEvent_t v = e->ev_pool.back();
        e->ev_pool.pop_back();
        return v;
    }
    hipEvent_t v = nullptr;
    if (hipEventCreate(&v) != hipSuccess)
        return nullptr;
    return v;
}

static int collect_timing(Engine* e)
{
    if (e->pending.empty())
        return FEDNL_OK;
    HIP_OK(hipStreamSynchronize(e->st));
    for (auto& p : e->pending) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, p.a, p.b));
        e->kms[p.kid] += ms;
        e->klaunch[p.kid] += 1;
        e->ev_pool.push_back(p.a);
        e->ev_pool.push_back(p.b);
    }
    e->pending.clear();
    return FEDNL_OK;
}

struct KTime {
    Engine* e;
    int kid;
    hipEvent_t b = nullptr;
    KTime(Engine* e_, int kid_) : e(e_), kid(kid_)
    {
        if (e->timing) {
            hipEvent_t a = pool_event(e);
            b = pool_event(e);
            hipEventRecord(a, e->st);
            e->pending.push_back({kid, a, b});
        }
    }
    ~KTime()
    {
        if (b)
            hipEventRecord(b, e->st);
    }
};

static int ensure_trace(Engine* e, int need)
{
    if (need <= e->trace_cap)
        return FEDNL_OK;
    int cap = std::max(need, std::max(1024, e->trace_cap * 2));
    double* nt = nullptr;
    HIP_OK(hipMalloc(&nt, sizeof(double) * 4 * (size_t)cap));
    HIP_OK(hipMemsetAsync(nt, 0, sizeof(double) * 4 * (size_t)cap, e->st));
    if (e->trace) {
        HIP_OK(hipMemcpyAsync(nt, e->trace, sizeof(double) * 4 * (size_t)e->trace_cap, hipMemcpyDeviceToDevice, e->st));
        HIP_OK(hipStreamSynchronize(e->st));
        HIP_OK(hipFree(e->trace));
    }
    e->trace = nt;
    e->trace_cap = cap;
    return FEDNL_OK;
}

static int launch_round(Engine* e)
{
    const int n = e->n;
    {
        KTime kt(e, KID_MARGIN);
        dim3 grid(e->m_pad / 16, n);
        hipLaunchKernelGGL(k_margin, grid, dim3(256), sizeof(double) * e->d_pad, e->st, e->A, e->x, e->m, e->m_pad,
                           e->d, e->d_pad, 1.0 / (double)e->m, e->Dw, e->gw, e->nlog);
    }
    {
        KTime kt(e, KID_HESS);
        const int nblk = 8 * ((n + 7) / 8) * e->tiles_up;
        hipLaunchKernelGGL(k_hessian_f64<false>, dim3(nblk), dim3(256), 0, e->st, e->A, e->Dw, e->gw, e->x, n,
                           e->m_pad, e->d, e->d_pad, e->tiles_up, e->o.lambda, e->S, e->learned, e->delta,
                           e->grad_c, e->part);
    }
    const int comp = e->o.compressor;
    {
        KTime kt(e, KID_SELECT);
        if (comp == FEDNL_TOPK) {
            hipLaunchKernelGGL(k_select_heap<false>, dim3(n), dim3(256), e->lds_select, e->st, e->delta, e->S, e->k,
                               e->part, e->tiles_up, e->lek_delta, (double*)nullptr, e->sel, e->sel_cnt, e->lek_p, e->lek_hit,
                               e->lek_popped);
        } else if (comp == FEDNL_TOPLEK) {
            hipLaunchKernelGGL(k_select_heap<true>, dim3(n), dim3(256), e->lds_select, e->st, e->delta, e->S, e->k,
                               e->part, e->tiles_up, e->lek_delta, (double*)nullptr, e->sel, e->sel_cnt, e->lek_p, e->lek_hit,
                               e->lek_popped);
        } else {
            if (comp == FEDNL_RANDK)
                hipLaunchKernelGGL(k_select_randk, dim3(n), dim3(256), e->lds_select, e->st, e->S, e->k, e->rng,
                                   e->sel, e->sel_cnt);
            else
                hipLaunchKernelGGL(k_select_window, dim3(n), dim3(256), 0, e->st, comp, e->S, e->k, e->rng, e->sel,
                                   e->sel_cnt, e->seq_start);
        }
    }
    {
        KTime kt(e, KID_EMIT);
        hipLaunchKernelGGL(k_emit, dim3(n), dim3(256), 0, e->st, comp, e->delta, e->S, e->k, e->lda, e->sel,
                           e->sel_cnt, e->lek_p, e->lek_hit, e->lek_popped, e->lek_state, e->workers, e->alpha,
                           e->prescale, e->mult, e->learned, e->pos, e->idx, e->vals, e->cnt);
    }
    {
        KTime kt(e, KID_MASTER_H);
        const int nblk = (int)((e->S + kWin - 1) / kWin);
        hipLaunchKernelGGL(k_master_h, dim3(nblk), dim3(256), 0, e->st, e->H, e->S, n, e->k, e->pos, e->vals, e->cnt);
    }
    {
        KTime kt(e, KID_MASTER_VEC);
        hipLaunchKernelGGL(k_master_vec, dim3(1), dim3(256), 0, e->st, comp, n, e->d, e->d_pad, e->m, e->m_pad,
                           e->grad_c, e->part, e->tiles_up, e->lk, e->nlog, e->x, e->o.lambda, e->inv_n, e->o.option_b, e->o.tracking,
                           e->lek_hit, e->lek_state, e->workers, e->g_avg, e->shift, e->trace, e->round_ctr,
                           e->trace_cap);
    }
    {
        KTime kt(e, KID_NEWTON);
        hipLaunchKernelGGL(k_newton, dim3(1), dim3(1024), e->lds_newton, e->st, e->d, e->dp, e->H, e->shift,
                           e->g_avg, e->W, e->x, e->dx, e->trace, e->round_ctr, e->trace_cap, 1);
    }
    return FEDNL_OK;
}

static size_t select_lds_bytes(int k)
{
    const int P = pow2_at_least(k);
    const size_t heap = (size_t)8 * k + (size_t)4 * round_up(P, 2) + (size_t)16 * kSeg;
    const size_t rk = (size_t)8 * P + (size_t)8 * k;
    return std::max(heap, rk);
}

static void free_engine(Engine* e)
{
    if (!e)
        return;
    void* ptrs[] = {e->A, e->x, e->Dw, e->gw, e->nlog, e->learned, e->delta, e->grad_c, e->part, e->lk,
                    e->sel, e->pos, e->idx, e->rng, e->lek_state, e->lek_popped, e->seq_start, e->sel_cnt,
                    e->cnt, e->lek_hit, e->round_ctr, e->lek_p, e->vals, e->H, e->g_avg, e->shift, e->W,
                    e->dx, e->trace};
    for (void* p : ptrs)
        if (p)
            hipFree(p);
    for (auto& p : e->pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (auto v : e->ev_pool) hipEventDestroy(v);
    for (auto v : e->round_ev) hipEventDestroy(v);
    if (e->ev_start)
        hipEventDestroy(e->ev_start);
    if (e->st)
        hipStreamDestroy(e->st);
    delete e;
}

template <class T>
static hipError_t dalloc(T** p, size_t count)
{
    return hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
}

static double theory_alpha(int comp, int alpha_mode, double alpha, int k, int d, size_t S)
{
    // TrainLoop.h:409-518; SpecialMathRoutinesForMatrix.h:195-218
    if (alpha_mode == 0)
        return alpha;
    if (comp == FEDNL_IDENTICAL)
        return 1.0 / (0.0 + 1.0);
    if (comp == FEDNL_NATURAL)
        return 1.0 / (1.0 / 8.0 + 1.0);
    if (comp == FEDNL_RANDK || comp == FEDNL_RANDSEQK) {
        const double w = ((double)S / k - 1.0);
        return 1.0 / (w + 1.0);
    }
    const double delta = (double)k / ((double)d * (double)d);  // DOPT_FIX_TOPK_CONTRACTION_FACTOR
    if (alpha_mode == 1)
        return 1.0 - sqrt(1.0 - delta);
    return 1.0;
}

static int create_engine(const fednl_opts* opts, int n, int m, int d, const double* features, const double* labels,
                         Engine** out)
{
    if (n <= 0 || m <= 0 || d <= 0 || !features || !labels) {
        set_err("fednl_create: empty problem");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (n > kMaxClients) {
        set_err("fednl_create: more than 1024 clients");
        return FEDNL_E_UNSUPPORTED;
    }
    if (d > kMaxNewtonD) {
        set_err("fednl_create: d > 512 needs the multi-workgroup Newton solve");
        return FEDNL_E_UNSUPPORTED;
    }
    Engine* e = new Engine();
    e->o = *opts;
    e->n = n;
    e->m = m;
    e->d = d;
    e->lda = round_up(d, 4);
    e->S = (size_t)d * (d + 1) / 2;
    const int comp = opts->compressor;
    if (comp == FEDNL_IDENTICAL)
        e->k = (int)e->S;  // TrainLoop.h:354-357
    else
        e->k = (int)(opts->k_mult * d);  // TrainLoop.h:362
    if (e->k <= 0 || (size_t)e->k > e->S) {
        set_err("fednl_create: k must be in [1, d(d+1)/2]");
        delete e;
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (comp != FEDNL_TOPK && comp != FEDNL_TOPLEK && comp != FEDNL_RANDK && comp != FEDNL_RANDSEQK &&
        comp != FEDNL_IDENTICAL) {
        set_err("fednl_create: compressor not supported by this engine");
        delete e;
        return FEDNL_E_UNSUPPORTED;
    }
    if ((comp == FEDNL_TOPK || comp == FEDNL_TOPLEK || comp == FEDNL_RANDK) && e->k > kMaxHeapK &&
        (size_t)e->k < e->S) {
        set_err("fednl_create: k above the LDS-resident selection limit (8192)");
        delete e;
        return FEDNL_E_UNSUPPORTED;
    }
    e->m_pad = round_up(m, kKC);
    e->d_pad = round_up(d, kTile);
    e->nt = e->d_pad / kTile;
    e->tiles_up = e->nt * (e->nt + 1) / 2;
    e->dp = round_up(d, 4);
    e->workers = std::max(1, std::min(opts->workers, n));
    e->alpha = theory_alpha(comp, opts->alpha_mode, opts->alpha, e->k, d, e->S);
    e->inv_n = 1.0 / (double)n;
    e->prescale = e->inv_n * e->alpha;
    e->mult = (double)e->S / (double)e->k;
    e->lek_delta = (double)e->k * (1.0 / (double)e->S);
    e->lds_select = select_lds_bytes(std::min<int>(e->k, kMaxHeapK));
    e->lds_newton = sizeof(double) * (size_t)kNB * e->dp;

    int rc = FEDNL_OK;
#define ALLOC(ptr, cnt)                                          \
    do {                                                         \
        hipError_t _e = dalloc(&(ptr), (cnt));                   \
        if (_e != hipSuccess) {                                  \
            set_err("hipMalloc " #ptr, _e);                      \
            free_engine(e);                                      \
            return FEDNL_E_DEVICE;                               \
        }                                                        \
    } while (0)
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate");
        delete e;
        return FEDNL_E_DEVICE;
    }
    const size_t nA = (size_t)n * e->m_pad * e->d_pad;
    ALLOC(e->A, nA);
    ALLOC(e->x, e->d_pad);
    ALLOC(e->Dw, (size_t)n * e->m_pad);
    ALLOC(e->gw, (size_t)n * e->m_pad);
    ALLOC(e->nlog, (size_t)n * e->m_pad);
    ALLOC(e->learned, (size_t)n * e->S);
    ALLOC(e->delta, (size_t)n * e->S);
    ALLOC(e->grad_c, (size_t)n * e->d_pad);
    ALLOC(e->part, (size_t)n * e->tiles_up * 2);
    ALLOC(e->lk, n);
    ALLOC(e->sel, (size_t)n * e->k);
    ALLOC(e->pos, (size_t)n * e->k);
    ALLOC(e->idx, (size_t)n * e->k);
    ALLOC(e->vals, (size_t)n * e->k);
    ALLOC(e->rng, n);
    ALLOC(e->lek_state, e->workers);
    ALLOC(e->lek_popped, n);
    ALLOC(e->seq_start, n);
    ALLOC(e->sel_cnt, n);
    ALLOC(e->cnt, n);
    ALLOC(e->lek_hit, n);
    ALLOC(e->lek_p, n);
    ALLOC(e->round_ctr, 1);
    ALLOC(e->H, e->S);
    ALLOC(e->g_avg, e->d_pad);
    ALLOC(e->shift, 1);
    ALLOC(e->W, (size_t)e->dp * e->dp);
    ALLOC(e->dx, e->d_pad);
#undef ALLOC
    if ((rc = ensure_trace(e, 1024)) != FEDNL_OK) {
        free_engine(e);
        return rc;
    }
    if (e->lds_select > 64 * 1024) {
        hipFuncSetAttribute((const void*)k_select_heap<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)e->lds_select);
        hipFuncSetAttribute((const void*)k_select_heap<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)e->lds_select);
        hipFuncSetAttribute((const void*)k_select_randk, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)e->lds_select);
    }
    if (e->lds_newton > 64 * 1024)
        hipFuncSetAttribute((const void*)k_newton, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->lds_newton);

    // host staging of the sign-folded, padded design
    {
        std::vector<double> hA((size_t)e->m_pad * e->d_pad);
        for (int c = 0; c < n; ++c) {
            std::fill(hA.begin(), hA.end(), 0.0);
            for (int s = 0; s < m; ++s) {
                const double bsgn = labels[(size_t)c * m + s];
                const double* src = features + ((size_t)c * m + s) * d;
                double* dst = hA.data() + (size_t)s * e->d_pad;
                for (int i = 0; i < d; ++i)
                    dst[i] = (bsgn == 1.0) ? src[i] : src[i] * -1.0;
            }
            if (hipMemcpy(e->A + (size_t)c * e->m_pad * e->d_pad, hA.data(), sizeof(double) * hA.size(),
                          hipMemcpyHostToDevice) != hipSuccess) {
                set_err("hipMemcpy A");
                free_engine(e);
                return FEDNL_E_DEVICE;
            }
        }
    }
    std::vector<uint32_t> seeds(n);
    for (int c = 0; c < n; ++c) seeds[c] = (uint32_t)(123 + c);  // TrainLoop.h:632
    std::vector<uint32_t> lek(e->workers, 1234u);                   // RandomGenRealLinear default seed
    hipMemcpyAsync(e->rng, seeds.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, e->st);
    hipMemcpyAsync(e->lek_state, lek.data(), sizeof(uint32_t) * e->workers, hipMemcpyHostToDevice, e->st);
    hipMemsetAsync(e->x, 0, sizeof(double) * e->d_pad, e->st);
    hipMemsetAsync(e->round_ctr, 0, sizeof(int), e->st);
    hipMemsetAsync(e->lek_hit, 0, sizeof(int) * n, e->st);
    hipMemsetAsync(e->grad_c, 0, sizeof(double) * (size_t)n * e->d_pad, e->st);

    // learned_c = H_c(x0), H = mean_c learned_c
    {
        dim3 grid(e->m_pad / 16, n);
        hipLaunchKernelGGL(k_margin, grid, dim3(256), sizeof(double) * e->d_pad, e->st, e->A, e->x, e->m, e->m_pad,
                           e->d, e->d_pad, 1.0 / (double)e->m, e->Dw, e->gw, e->nlog);
        const int nblk = 8 * ((n + 7) / 8) * e->tiles_up;
        hipLaunchKernelGGL(k_hessian_f64<true>, dim3(nblk), dim3(256), 0, e->st, e->A, e->Dw, e->gw, e->x, n,
                           e->m_pad, e->d, e->d_pad, e->tiles_up, e->o.lambda, e->S, e->learned, e->delta,
                           e->grad_c, e->part);
        hipLaunchKernelGGL(k_init_H, dim3((unsigned)((e->S + 255) / 256)), dim3(256), 0, e->st, e->H, e->learned,
                           e->S, n, e->inv_n);
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        set_err("kernel launch (create)", le);
        free_engine(e);
        return FEDNL_E_DEVICE;
    }
    if (hipStreamSynchronize(e->st) != hipSuccess) {
        set_err("hipStreamSynchronize (create)");
        free_engine(e);
        return FEDNL_E_DEVICE;
    }
    *out = e;
    return FEDNL_OK;
}

static int enqueue_rounds(Engine* e, int rounds)
{
    if (rounds <= 0)
        return FEDNL_OK;
    int rc = ensure_trace(e, e->rounds_enqueued + rounds + 1);
    if (rc != FEDNL_OK)
        return rc;
    if (!e->ev_start) {
        HIP_OK(hipEventCreate(&e->ev_start));
        HIP_OK(hipEventRecord(e->ev_start, e->st));
    }
    for (int r = 0; r < rounds; ++r) {
        launch_round(e);
        hipEvent_t ev = nullptr;
        HIP_OK(hipEventCreate(&ev));
        HIP_OK(hipEventRecord(ev, e->st));
        e->round_ev.push_back(ev);
        e->rounds_enqueued += 1;
    }
    HIP_OK(hipGetLastError());
    return FEDNL_OK;
}

}  // namespace fednl

using namespace fednl;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void fednl_default_opts(fednl_opts* o)
{
    memset(o, 0, sizeof(*o));
    o->lambda = 1e-3;
    o->compressor = FEDNL_TOPK;
    o->k_mult = 8.0;
    o->alpha_mode = 2;
    o->alpha = 1.0;
    o->option_b = 1;
    o->workers = 1;
    o->tracking = 1;
    o->hessian_mode = 0;
}

const char* fednl_last_error(void) { return g_err.c_str(); }

int fednl_create(const fednl_opts* opts, int n, int m, int d, const double* features, const double* labels,
                 void** handle)
{
    if (!opts || !handle) {
        set_err("fednl_create: null argument");
        return FEDNL_E_MISSING_ARGUMENT;
    }
    Engine* e = nullptr;
    int rc = create_engine(opts, n, m, d, features, labels, &e);
    *handle = (rc == FEDNL_OK) ? (void*)e : nullptr;
    return rc;
}

void fednl_destroy(void* handle) { free_engine((Engine*)handle); }

int fednl_sizes(void* handle, int* k, int64_t* S, int* lda)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    if (k) *k = e->k;
    if (S) *S = (int64_t)e->S;
    if (lda) *lda = e->lda;
    return FEDNL_OK;
}

double fednl_alpha(void* handle) { return handle ? ((Engine*)handle)->alpha : 0.0; }

int fednl_enqueue(void* handle, int rounds)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    return enqueue_rounds(e, rounds);
}

int fednl_sync(void* handle)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    return collect_timing(e);
}

int fednl_run(void* handle, int rounds)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    const int first = e->rounds_enqueued;
    int rc = enqueue_rounds(e, rounds);
    if (rc != FEDNL_OK)
        return rc;
    rc = fednl_sync(handle);
    if (rc != FEDNL_OK)
        return rc;
    std::vector<double> tr((size_t)4 * rounds);
    HIP_OK(hipMemcpy(tr.data(), e->trace + (size_t)4 * first, sizeof(double) * tr.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < rounds; ++r)
        if (tr[(size_t)4 * r + 3] != 1.0) {
            set_err("Cholesky factorization failed (non-positive pivot)");
            return FEDNL_E_INTERNAL;
        }
    return FEDNL_OK;
}

int fednl_trace(void* handle, double* grad_norm, double* fval, double* x_norm, double* seconds, int* chol_ok,
                int* count)
{
    Engine* e = (Engine*)handle;
    if (!e || !count)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    const int R = e->rounds_enqueued;
    *count = R;
    if (R == 0)
        return FEDNL_OK;
    std::vector<double> tr((size_t)4 * R);
    HIP_OK(hipMemcpy(tr.data(), e->trace, sizeof(double) * tr.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < R; ++r) {
        if (grad_norm) grad_norm[r] = tr[(size_t)4 * r + 0];
        if (fval) fval[r] = tr[(size_t)4 * r + 1];
        if (x_norm) x_norm[r] = tr[(size_t)4 * r + 2];
        if (chol_ok) chol_ok[r] = tr[(size_t)4 * r + 3] == 1.0 ? 1 : 0;
        if (seconds) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e->ev_start, e->round_ev[r]) != hipSuccess)
                ms = -1.f;
            seconds[r] = ms * 1e-3;
        }
    }
    return FEDNL_OK;
}

int fednl_get_x(void* handle, double* x)
{
    Engine* e = (Engine*)handle;
    if (!e || !x)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    HIP_OK(hipMemcpy(x, e->x, sizeof(double) * e->d, hipMemcpyDeviceToHost));
    return FEDNL_OK;
}

int fednl_last_round(void* handle, double* grad, double* lk, int* cnt, uint32_t* pos, double* vals, uint32_t* flat_idx,
                     double* dx)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    if (grad) HIP_OK(hipMemcpy(grad, e->g_avg, sizeof(double) * e->d, hipMemcpyDeviceToHost));
    if (lk) HIP_OK(hipMemcpy(lk, e->lk, sizeof(double) * e->n, hipMemcpyDeviceToHost));
    if (cnt) HIP_OK(hipMemcpy(cnt, e->cnt, sizeof(int) * e->n, hipMemcpyDeviceToHost));
    if (pos) HIP_OK(hipMemcpy(pos, e->pos, sizeof(uint32_t) * (size_t)e->n * e->k, hipMemcpyDeviceToHost));
    if (vals) HIP_OK(hipMemcpy(vals, e->vals, sizeof(double) * (size_t)e->n * e->k, hipMemcpyDeviceToHost));
    if (flat_idx) HIP_OK(hipMemcpy(flat_idx, e->idx, sizeof(uint32_t) * (size_t)e->n * e->k, hipMemcpyDeviceToHost));
    if (dx) HIP_OK(hipMemcpy(dx, e->dx, sizeof(double) * e->d, hipMemcpyDeviceToHost));
    return FEDNL_OK;
}

int fednl_get_state(void* handle, double* x, double* learned, double* H, uint32_t* rng, uint32_t* toplek_rng)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    if (x) HIP_OK(hipMemcpy(x, e->x, sizeof(double) * e->d, hipMemcpyDeviceToHost));
    if (learned) HIP_OK(hipMemcpy(learned, e->learned, sizeof(double) * (size_t)e->n * e->S, hipMemcpyDeviceToHost));
    if (H) HIP_OK(hipMemcpy(H, e->H, sizeof(double) * e->S, hipMemcpyDeviceToHost));
    if (rng) HIP_OK(hipMemcpy(rng, e->rng, sizeof(uint32_t) * e->n, hipMemcpyDeviceToHost));
    if (toplek_rng) HIP_OK(hipMemcpy(toplek_rng, e->lek_state, sizeof(uint32_t) * e->workers, hipMemcpyDeviceToHost));
    return FEDNL_OK;
}

int fednl_set_state(void* handle, const double* x, const double* learned, const double* H, const uint32_t* rng,
                    const uint32_t* toplek_rng)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    if (x) HIP_OK(hipMemcpy(e->x, x, sizeof(double) * e->d, hipMemcpyHostToDevice));
    if (learned) HIP_OK(hipMemcpy(e->learned, learned, sizeof(double) * (size_t)e->n * e->S, hipMemcpyHostToDevice));
    if (H) HIP_OK(hipMemcpy(e->H, H, sizeof(double) * e->S, hipMemcpyHostToDevice));
    if (rng) HIP_OK(hipMemcpy(e->rng, rng, sizeof(uint32_t) * e->n, hipMemcpyHostToDevice));
    if (toplek_rng) HIP_OK(hipMemcpy(e->lek_state, toplek_rng, sizeof(uint32_t) * e->workers, hipMemcpyHostToDevice));
    return FEDNL_OK;
}

int fednl_kernel_timing(void* handle, int enable)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    int rc = collect_timing(e);
    e->timing = enable != 0;
    for (int q = 0; q < KID_COUNT; ++q) {
        e->kms[q] = 0.0;
        e->klaunch[q] = 0;
    }
    return rc;
}

int fednl_kernel_times(void* handle, double* ms, int* launches, int max_kernels, const char** names)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    int rc = collect_timing(e);
    if (rc != FEDNL_OK)
        return rc;
    const int nk = std::min(max_kernels, (int)KID_COUNT);
    for (int q = 0; q < nk; ++q) {
        if (ms) ms[q] = e->kms[q];
        if (launches) launches[q] = e->klaunch[q];
        if (names) names[q] = kKernelNames[q];
    }
    return KID_COUNT;
}

int fednl_select(int compressor, int nclients, int d, int k, const double* delta, uint32_t* rng, uint32_t* toplek_rng,
                 int workers, double delta_param, uint32_t* out_pos, int* out_cnt)
{
    if (nclients <= 0 || d <= 0 || k <= 0 || !delta || !out_pos || !out_cnt)
        return FEDNL_E_MISSING_ARGUMENT;
    const size_t S = (size_t)d * (d + 1) / 2;
    if ((size_t)k > S)
        return FEDNL_E_WRONG_ARGUMENT;
    if (k > kMaxHeapK && (size_t)k < S) {
        set_err("fednl_select: k above the LDS-resident selection limit");
        return FEDNL_E_UNSUPPORTED;
    }
    workers = std::max(1, std::min(workers, nclients));
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double *ddelta = nullptr, *learned = nullptr, *lek_p = nullptr, *vals = nullptr;
    uint32_t *drng = nullptr, *dlek = nullptr, *sel = nullptr, *pos = nullptr, *idx = nullptr, *popped = nullptr;
    int *sel_cnt = nullptr, *hit = nullptr, *cnt = nullptr;
    int rc = FEDNL_OK;
    auto cleanup = [&]() {
        void* ps[] = {ddelta, learned, lek_p, vals, drng, dlek, sel, pos, idx, popped, sel_cnt, hit, cnt};
        for (void* p : ps)
            if (p) hipFree(p);
        hipStreamDestroy(st);
    };
#define SALLOC(p, count)                                   \
    if (dalloc(&(p), (count)) != hipSuccess) {             \
        set_err("hipMalloc (select)");                     \
        cleanup();                                         \
        return FEDNL_E_DEVICE;                             \
    }
    SALLOC(ddelta, (size_t)nclients * S);
    SALLOC(learned, (size_t)nclients * S);
    SALLOC(lek_p, nclients);
    SALLOC(vals, (size_t)nclients * k);
    SALLOC(drng, nclients);
    SALLOC(dlek, workers);
    SALLOC(sel, (size_t)nclients * k);
    SALLOC(pos, (size_t)nclients * k);
    SALLOC(idx, (size_t)nclients * k);
    SALLOC(popped, nclients);
    SALLOC(sel_cnt, nclients);
    SALLOC(hit, nclients);
    SALLOC(cnt, nclients);
#undef SALLOC
    hipMemcpyAsync(ddelta, delta, sizeof(double) * (size_t)nclients * S, hipMemcpyHostToDevice, st);
    hipMemsetAsync(learned, 0, sizeof(double) * (size_t)nclients * S, st);
    hipMemsetAsync(hit, 0, sizeof(int) * nclients, st);
    if (rng) hipMemcpyAsync(drng, rng, sizeof(uint32_t) * nclients, hipMemcpyHostToDevice, st);
    if (toplek_rng) hipMemcpyAsync(dlek, toplek_rng, sizeof(uint32_t) * workers, hipMemcpyHostToDevice, st);
    const size_t lds = select_lds_bytes(std::min(k, kMaxHeapK));
    if (lds > 64 * 1024) {
        hipFuncSetAttribute((const void*)k_select_heap<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipFuncSetAttribute((const void*)k_select_heap<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipFuncSetAttribute((const void*)k_select_randk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    }
    // TopLEK needs ||Delta||^2: computed on the host-side copy here as tile partials would be
    // in the round; the standalone path passes one partial per client.
    double* parts = nullptr;
    if (compressor == FEDNL_TOPLEK) {
        // ||Delta||^2 in the summation order of dopt's AVX2 build
        // (MatrixNMD_SIMD_double.h:1067-1156): the partial pair (off, diag) is
        // laid out so that (off + off) + diag reproduces its result.
        std::vector<double> hp((size_t)nclients * 2);
        for (int c = 0; c < nclients; ++c) {
            const double* up = delta + (size_t)c * S;
            double acc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            double rest = 0.0, dg = 0.0;
            size_t col = 0;
            for (int j = 0; j < d; ++j) {
                const double* a = up + col;
                const size_t items = (size_t)j / 8 * 8;
                size_t i = 0;
                for (; i < items; i += 8)
                    for (int u = 0; u < 2; ++u)
                        for (int l = 0; l < 4; ++l) {
                            const double v = a[i + 4 * u + l];
                            const double sq = v * v;
                            acc[u][l] = acc[u][l] + sq;
                        }
                for (; i < (size_t)j; ++i) {
                    const double v = a[i];
                    const double sq = v * v;
                    rest = rest + sq;
                }
                const double v = a[j];
                const double sq = v * v;
                dg = dg + sq;
                col += (size_t)j + 1;
            }
            double out = 0.0;
            for (int u = 0; u < 2; ++u) {
                const double lo = acc[u][0] + acc[u][2];
                const double hi = acc[u][1] + acc[u][3];
                const double h = lo + hi;
                out = out + h;
            }
            out = out + rest;
            hp[(size_t)c * 2] = out;
            hp[(size_t)c * 2 + 1] = dg;
        }
        if (dalloc(&parts, hp.size()) != hipSuccess) {
            cleanup();
            return FEDNL_E_DEVICE;
        }
        hipMemcpyAsync(parts, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice, st);
        hipStreamSynchronize(st);
    }
    if (compressor == FEDNL_TOPK)
        hipLaunchKernelGGL(k_select_heap<false>, dim3(nclients), dim3(256), lds, st, ddelta, S, k, (const double*)nullptr,
                           0, delta_param, (double*)nullptr, sel, sel_cnt, lek_p, hit, popped);
    else if (compressor == FEDNL_TOPLEK)
        hipLaunchKernelGGL(k_select_heap<true>, dim3(nclients), dim3(256), lds, st, ddelta, S, k, parts, 1,
                           delta_param, (double*)nullptr, sel, sel_cnt, lek_p, hit, popped);
    else if (compressor == FEDNL_RANDK)
        hipLaunchKernelGGL(k_select_randk, dim3(nclients), dim3(256), lds, st, S, k, drng, sel, sel_cnt);
    else if (compressor == FEDNL_RANDSEQK || compressor == FEDNL_IDENTICAL)
        hipLaunchKernelGGL(k_select_window, dim3(nclients), dim3(256), 0, st, compressor, S, k, drng, sel, sel_cnt,
                           (uint32_t*)nullptr);
    else {
        cleanup();
        return FEDNL_E_UNSUPPORTED;
    }
    hipLaunchKernelGGL(k_emit, dim3(nclients), dim3(256), 0, st, compressor, ddelta, S, k, round_up(d, 4), sel, sel_cnt,
                       lek_p, hit, popped, dlek, workers, 1.0, 1.0, 1.0, learned, pos, idx, vals, cnt);
    if (compressor == FEDNL_TOPLEK) {
        // advance the streams as the master kernel does
        std::vector<int> hh(nclients);
        hipMemcpyAsync(hh.data(), hit, sizeof(int) * nclients, hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
        for (int w = 0; w < workers && toplek_rng; ++w) {
            uint32_t s = toplek_rng[w];
            for (int c = w; c < nclients; c += workers)
                if (hh[c]) s = (kLcgA * s + kLcgC) % 4294967295u;
            toplek_rng[w] = s;
        }
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        set_err("kernel launch (select)", le);
        rc = FEDNL_E_DEVICE;
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        set_err("hipStreamSynchronize (select)");
        rc = FEDNL_E_DEVICE;
    }
    if (rc == FEDNL_OK) {
        hipMemcpy(out_pos, pos, sizeof(uint32_t) * (size_t)nclients * k, hipMemcpyDeviceToHost);
        hipMemcpy(out_cnt, cnt, sizeof(int) * nclients, hipMemcpyDeviceToHost);
        if (rng) hipMemcpy(rng, drng, sizeof(uint32_t) * nclients, hipMemcpyDeviceToHost);
    }
    if (parts) hipFree(parts);
    cleanup();
    return rc;
}

int fednl_probe_mfma_f64(double* out768)
{
    if (!out768)
        return FEDNL_E_MISSING_ARGUMENT;
    double* d_out = nullptr;
    HIP_OK(dalloc(&d_out, 768));
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, d_out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(out768, d_out, sizeof(double) * 768, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(d_out));
    return FEDNL_OK;
}

int fednl_newton_solve(int d, const double* H_up, double shift, const double* b, double* dx)
{
    if (d <= 0 || !H_up || !b || !dx)
        return FEDNL_E_MISSING_ARGUMENT;
    if (d > kMaxNewtonD) {
        set_err("fednl_newton_solve: d > 512");
        return FEDNL_E_UNSUPPORTED;
    }
    const size_t S = (size_t)d * (d + 1) / 2;
    const int dp = round_up(d, 4);
    double *dH = nullptr, *dsh = nullptr, *db = nullptr, *dW = nullptr, *dx_d = nullptr, *dtr = nullptr;
    int* dctr = nullptr;
    HIP_OK(dalloc(&dH, S));
    HIP_OK(dalloc(&dsh, 1));
    HIP_OK(dalloc(&db, d));
    HIP_OK(dalloc(&dW, (size_t)dp * dp));
    HIP_OK(dalloc(&dx_d, d));
    HIP_OK(dalloc(&dtr, 4));
    HIP_OK(dalloc(&dctr, 1));
    HIP_OK(hipMemcpy(dH, H_up, sizeof(double) * S, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dsh, &shift, sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(db, b, sizeof(double) * d, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(dctr, 0, sizeof(int)));
    HIP_OK(hipMemset(dtr, 0, sizeof(double) * 4));
    const size_t lds = sizeof(double) * (size_t)kNB * dp;
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void*)k_newton, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_newton, dim3(1), dim3(1024), lds, 0, d, dp, dH, dsh, db, dW, (double*)nullptr, dx_d, dtr, dctr,
                       1, 0);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    double tr[4];
    HIP_OK(hipMemcpy(tr, dtr, sizeof(tr), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(dx, dx_d, sizeof(double) * d, hipMemcpyDeviceToHost));
    hipFree(dH); hipFree(dsh); hipFree(db); hipFree(dW); hipFree(dx_d); hipFree(dtr); hipFree(dctr);
    return tr[3] == 1.0 ? 1 : 0;
}

}  // extern "C"
