// fednl_kernels.hip -- MI355X (gfx950) kernels and host engine of the FedNL
// round. One FedNL round of dopt's single-node simulation
// (dopt/bin_fednl_local/include/TrainLoop.h:1054-1508 with the client loops of
// LocalWorkerForFedNL_{TopK,TopLEK,RandK,RandSeqK}.h) runs as five kernels on
// one stream (six with TopLEK), every client simulated in place:
//
//   k_margin        t = A x, sigma, D = (1-sigma)sigma/m, g = (1-sigma)(-1/m), -log sigma
//   k_hessian_f64   upper tiles of A^T diag(D) A on v_mfma_f64_16x16x4f64, fused with
//                   +lambda, Delta = H - learned, Lk partial sums, and the client gradient
//   k_select_*      index selection (heap scan of TopK/TopLEK, RandK shuffle, RandSeqK
//                   window), followed in the same workgroup by the client's message
//                   values and learned-Hessian update (k_emit on its own for TopLEK,
//                   whose random rounding runs over the other clients)
//   k_master        H[idx] += val in client order (deterministic), gradient / Lk / f
//                   averages, the shift, the trace record
//   k_newton_mc     Cholesky of H + shift I, two triangular solves, x -= dx, over
//                   one workgroup per 32-column block (left-looking, flag handoffs)
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
#include <atomic>
#include <chrono>
#include <mutex>
#include <new>
#include <queue>
#include <string>
#include <type_traits>
#include <vector>
#include <thread>

#include "fednl_hip.h"

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

namespace fednl {

// Device-clock probes for phase timing in a diagnostic build (-DFEDNL_PROBE);
// they compile to nothing otherwise.
#ifdef FEDNL_PROBE
__device__ unsigned long long g_probe[16384];
#define PROBE_T(var) const unsigned long long var = wall_clock64()
#define PROBE_ADD(slot, t0) \
    do { if (threadIdx.x == 0) atomicAdd(&g_probe[(slot)], wall_clock64() - (t0)); } while (0)
#define PROBE_SET(slot, v) \
    do { if (threadIdx.x == 0) g_probe[(slot)] = (unsigned long long)(v); } while (0)
#define PROBE_ADD_AT(slot, t0, thread) \
    do { if (threadIdx.x == (thread)) atomicAdd(&g_probe[(slot)], wall_clock64() - (t0)); } while (0)
#define PROBE_STAMP_AT(slot, thread) \
    do { if (threadIdx.x == (thread)) g_probe[(slot)] = wall_clock64(); } while (0)
#else
#define PROBE_ADD_AT(slot, t0, thread) do {} while (0)
#define PROBE_STAMP_AT(slot, thread) do {} while (0)
#define PROBE_T(var) do {} while (0)
#define PROBE_ADD(slot, t0) do {} while (0)
#define PROBE_SET(slot, v) do {} while (0)
#endif

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
constexpr int kMaxHeapK = 8192;    // LDS-resident heap limit
constexpr int kMaxSelK = 32768;    // TopK / TopLEK / RandK: largest k (the tail sorts k ids in LDS)
constexpr int kEmitLdsMax = 128 * 1024;  // largest selection bitmap kept in LDS (bytes)
constexpr int kScanGroup = 16;     // TopK scan: chunks of 64 positions per staged group
constexpr int kMaxScanRing = 8;    // TopK scan: most LDS slots of staged groups
static_assert(kScanGroup % 2 == 0, "the scan tests chunk pairs");
constexpr int kMaxClients = 1024;  // master scatter client table
// per-round trace record: ||grad f||, f, ||x||, factorization flag,
// line-search trials, participating clients, Hessian items they sent
constexpr int kTraceStride = 9;
// per-round trace record: ||grad f||, f, ||x||, factorization flag,
// line-search trials, participating clients, items sent, then the device
// wall clock (ticks) when the round's first kernel started and when its last
// kernel finished
constexpr int kTraceT0 = 7, kTraceT1 = 8;
constexpr int kMaxLineSearchSteps = 64;  // trial launches enqueued per round
// master state (mst): line-search scalars, FedNL-PP's running Lk average,
// then x_k (d) from kLsHead on
enum MstSlot { LS_FXK, LS_T, LS_DD, LS_DONE, LS_TRIALS, PP_LI, kLsHead = 8 };
constexpr int kNB = 32;            // Cholesky panel width
constexpr int kMaxNewtonD = 512;   // largest d of the no-allocation Newton solve
constexpr int kMaxD = 2048;        // features per problem

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

// list position p is a diagonal entry (j, j) iff p = j(j+3)/2, i.e. 8p+9 = (2j+3)^2.
// The root's estimate comes from the hardware's single-precision square root
// (one instruction) and is corrected in integers, so the test is exact for
// every p (8p + 9 below 2^62) at a fraction of a double sqrt's cost.
__device__ inline bool pos_is_diag(uint64_t p)
{
    const uint64_t v = 8 * p + 9;
    uint64_t r = (uint64_t)__builtin_amdgcn_sqrtf((float)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r * r == v;
}

// dopt's Natural compressor on one value (MatrixNMD.h:2009-2029): the mantissa
// bits cleared, sign and exponent kept
__device__ inline double natural_round(double v)
{
    return __longlong_as_double(__double_as_longlong(v) & (long long)0xFFF0000000000000ull);
}

// the TopK tail's bitmap of S list positions fits in the heap's k + 4 values
__host__ __device__ inline bool select_bitmap_fits(size_t S, int k)
{
    return ((S + 31) / 32) * 4 <= (size_t)8 * (k + 4);
}

// The TopK scan's waves hand groups through LDS stamps: bounded spins (a
// wave that waits longer than any scan can take gives up rather than hang).
__device__ inline void scan_lds_wait_eq(int* f, int v)
{
    long n = 0;
    while (*(volatile int*)f != v && ++n < (1L << 24))
        __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ inline void scan_lds_wait_ge(int* f, int v)
{
    long n = 0;
    while (*(volatile int*)f < v && ++n < (1L << 24))
        __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Global-memory views of the descriptor's arrays. The pointers in ProblemDev
// are generic; accesses through them compile to flat instructions, which
// count against the LDS wait counter too, so every LDS wait would also wait
// for the flat loads in flight. Kernels that mix LDS work with global
// prefetches take their arrays through G() and access them with global_*.
// (The host pass parses the same device code; there the views are plain
// pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gmem = __attribute__((address_space(1))) T;
#else
template <class T>
using gmem = T;
#endif
template <class T>
__device__ __forceinline__ gmem<T>* G(T* p)
{
    return (gmem<T>*)p;
}
typedef gmem<const double> gdouble;

// value of a double held by lane `src` (uniform), without an LDS round trip
__device__ inline double bcast_lane(double v, int src)
{
    const unsigned long long u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(u & 0xFFFFFFFFull), src);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
    return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

// DPP move of a double (both halves with the same control: quad_perm
// patterns 0x00-0xFF, row_shl:n = 0x100 + n, ...)
template <int kPerm>
__device__ inline double dpp_quad(double v)
{
    const unsigned long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u & 0xFFFFFFFFull), kPerm, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), kPerm, 0xF, 0xF, false);
    return __longlong_as_double(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo);
}

// minimum over each aligned group of four lanes, in every lane of the group
__device__ inline double quad_min(double v)
{
    v = fmin(v, dpp_quad<0xB1>(v));  // lanes (1,0,3,2)
    v = fmin(v, dpp_quad<0x4E>(v));  // lanes (2,3,0,1)
    return v;
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
// Per-problem device descriptor. Every round kernel takes an array of these
// and a block map: entry b of the map is (problem, local block) for global
// block b, so one launch can serve many independent problems (a batch); a
// null map means one problem and local block = blockIdx.x. Map entries with
// problem < 0 are padding (the Hessian kernels align each problem's blocks to
// 8 so that blockIdx % 8 still names the XCD).
// ---------------------------------------------------------------------------
struct ProblemDev {
    int n, m, d, k, lda, m_pad, d_pad, tiles_up, dp, workers, comp, hmode, option_b, tracking, bm_words,
        trace_cap, update_x, ls, ls_max, pp, pp_tau;
    int nparts;         // Lk partial slots per client: Hessian tiles (dense), segments (sparse)
    size_t S;
    double lambda, alpha, prescale, inv_n, mult, lek_delta, ls_c, ls_gamma;
    double gd_step;     // GD: the global step size
    const double* A;
    const float* A32;
    const __bf16* AThi;
    const __bf16* ATlo;
    double *x, *Dw, *gw, *nlog, *learned, *delta, *grad_c, *part, *lk, *vals, *H, *g_avg, *shift, *W, *dx, *trace,
        *lek_p;
    double* mst;        // master state (MstSlot header, then x_k)
    // reduced-precision Hessian modes: per-16-row partial client gradients
    // from the margin pass (n x m_pad/16 x d_pad), summed by the Hessian
    // kernels' diagonal tiles; null: k_grad_f64 computes grad_c
    double* gpart;
    // sparse Hessian form (binary designs): per (client, dispatch rank) the
    // segment record (segment index, stream length, stream offset in 16-bit
    // tokens), the lane offsets inside its stream (kSpLanes + 1), the
    // streams, each sample's label sign
    const uint4* sp_seg;
    const uint32_t* sp_lane;
    const uint16_t* sp_stream;
    const double* sp_b;
    // and for its margin pass: per client m + 1 row offsets into the column
    // lists, each sample's nonzero columns (ascending)
    const uint32_t* sp_rowptr;
    const uint16_t* sp_cols;
    // FedNL-PP: selected flags, client sampler state, per-client direction
    // (n x d_pad) and Lk, this round's messages, H_c(x) (option b), full gradient
    int* pp_sel;
    uint32_t* pp_rng;
    double *pp_dir, *pp_lk, *pp_gmsg, *pp_lkmsg, *pp_h, *pp_tg;
    const double* H0;   // H after creation: the lower triangle of the master's full H
    const double* hup;  // Newton: packed H to fill W from (standalone solve), or null
    const double* nb;   // Newton: right-hand side
    uint32_t *sel, *pos, *idx, *rng, *lek_state, *lek_popped, *seq_start, *sel_bm;
    int *sel_cnt, *cnt, *lek_hit, *round_ctr;
    // Newton over many workgroups: sync words (epoch, ticket, failure epoch,
    // error, per block column: segments done, diagonal block done), the
    // forward solution y, the reciprocal pivots and each diagonal block's
    // L^-T (32 x 32 per block column)
    int* nsync;
    double *ny, *ninv, *nxt;
    int sel_ring;       // TopK / TopLEK scan: LDS slots of staged groups (2 .. kMaxScanRing)
    // wide selection (k above the LDS heap or S above the LDS bitmap): the
    // heap (or RandK's sort keys) in per-client global scratch, the emit
    // bitmap written straight to sel_bm
    int wide;
    unsigned char* sel_scr;
    size_t sel_scr_stride;  // bytes per client
};

#define FEDNL_BLOCK()                                                                     \
    const int2 bmv_ = bmap ? bmap[blockIdx.x] : make_int2(0, (int)blockIdx.x);            \
    if (bmv_.x < 0)                                                                       \
        return;                                                                           \
    const ProblemDev& PD = probs[bmv_.x];                                                 \
    const int blk = bmv_.y

#define FEDNL_KARGS const ProblemDev* __restrict__ probs, const int2* __restrict__ bmap

// FedNL-PP: client c sits this round out
__device__ inline bool pp_idle(const ProblemDev& P, int c) { return P.pp && P.pp_sel[c] == 0; }

// ---------------------------------------------------------------------------
// k_margin: kRows sample rows per wave. kTrial: a line-search trial point
// (only the -log sigma partials for f; nothing once the line search has
// stopped). kGrad (reduced-precision Hessian modes): the workgroup's 16 rows
// also give their part of the client gradient, sum_s A_c[s] g_s over the
// rows (logistic_regression.h:148-161), from the rows the margins were read
// from (held in registers while d_pad <= 64 kGradRegs), written to
// gpart[c][16-row block][d_pad]; the Hessian kernels' diagonal tiles add the
// blocks up, so A is read once per round in FP64. Dynamic LDS: x, then
// (kGrad) the four waves' column sums, 8 d_pad doubles each.
// ---------------------------------------------------------------------------
constexpr int kGradRegs = 8;
template <bool kTrial, int kRows, bool kGrad = false>
__global__ __launch_bounds__(256) void k_margin(FEDNL_KARGS)
{
    // kRows rows per wave, 4 kRows per workgroup (m_pad is a multiple of 32);
    // the -log sigma partials stay per 16 rows, whatever kRows is, so that f
    // is summed in the same order
    static_assert(kRows == 4 || kRows == 8, "16 or 32 rows per workgroup");
    static_assert(!kGrad || (kRows == 4 && !kTrial), "gradient partials per 16-row workgroup");
    constexpr int kRB = 4 * kRows;
    FEDNL_BLOCK();
    if (kTrial && (!PD.ls || PD.mst[LS_DONE] != 0.0))
        return;
    if (!kTrial && blk == 0 && threadIdx.x == 0 && PD.round_ctr) {
        // the round starts (the first kernel of every round)
        const int r = *PD.round_ctr;
        if (r < PD.trace_cap)
            PD.trace[(size_t)r * kTraceStride + kTraceT0] = (double)wall_clock64();
    }
    const gmem<double>* __restrict__ A = G(PD.A);
    const gmem<double>* __restrict__ x = G(PD.x);
    const int m = PD.m, m_pad = PD.m_pad, d = PD.d, d_pad = PD.d_pad;
    const double inv_m = 1.0 / (double)m;
    gmem<double>* __restrict__ Dw = G(PD.Dw);
    gmem<double>* __restrict__ gw = G(PD.gw);
    gmem<double>* __restrict__ nlog_part = G(PD.nlog);
    const int gxb = m_pad / kRB, bx = blk % gxb, c = blk / gxb;
    extern __shared__ double xs[];
    __shared__ double nl[kRB];
    if (threadIdx.x < kRB)
        nl[threadIdx.x] = 0.0;
    for (int i = threadIdx.x; i < d_pad; i += blockDim.x)
        xs[i] = (i < d) ? x[i] : 0.0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // the wave's rows, their loads interleaved; every lane's sum runs over the
    // same columns in the same order for each row
    const int s0 = bx * kRB + wave * kRows;
    const gmem<double>* row0 = A + ((size_t)c * m_pad + s0) * d_pad;
    double acc[kRows];
#pragma unroll
    for (int q = 0; q < kRows; ++q) acc[q] = 0.0;
    // kGrad: the rows' values stay in registers for the gradient part (the
    // same sums in the same order as the loop below)
    double av[kRows][kGrad ? kGradRegs : 1];
    bool in_regs = false;
    if constexpr (kGrad) {
        in_regs = d_pad <= 64 * kGradRegs;
        if (in_regs) {
#pragma unroll
            for (int j = 0; j < kGradRegs; ++j) {
                const int i = lane + 64 * j;
                if (i < d_pad) {
                    const double xi = xs[i];
#pragma unroll
                    for (int q = 0; q < kRows; ++q) {
                        av[q][j] = row0[(size_t)q * d_pad + i];
                        acc[q] = fma(av[q][j], xi, acc[q]);
                    }
                }
            }
        }
    }
    if (!in_regs) {
#pragma unroll 2
        for (int i = lane; i < d_pad; i += 64) {
            const double xi = xs[i];
#pragma unroll
            for (int q = 0; q < kRows; ++q)
                acc[q] = fma(row0[(size_t)q * d_pad + i], xi, acc[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < kRows; ++q)
        for (int off = 32; off > 0; off >>= 1)
            acc[q] += __shfl_xor(acc[q], off);
    // every lane holds the sums; lane q finishes row s0 + q
    double gq = 0.0;
    if (lane < kRows) {
        double t = acc[0];
#pragma unroll
        for (int q = 1; q < kRows; ++q)
            t = (lane == q) ? acc[q] : t;
        const int s = s0 + lane;
        const size_t o = (size_t)c * m_pad + s;
        if (s < m) {
            // logistic_regression.h:98-101,148-161,398-404
            const double sg = 1.0 / (1.0 + exp(-t));
            if (!kTrial) {
                const double z1 = 1.0 - sg;
                const double z2 = z1 * sg;
                Dw[o] = inv_m * z2;
                gq = z1 * (-inv_m);
                gw[o] = gq;
            }
            nl[wave * kRows + lane] = -log(sg);
        } else {
            if (!kTrial) {
                Dw[o] = 0.0;
                gw[o] = 0.0;
            }
            nl[wave * kRows + lane] = 0.0;
        }
    }
    if constexpr (kGrad) {
        // this workgroup's 16 rows: per column, each wave's four rows in
        // order into gl[wave]; after the barrier, the waves in the order
        // (0 + 1) + (2 + 3)
        double g[kRows];
#pragma unroll
        for (int q = 0; q < kRows; ++q) g[q] = __shfl(gq, q);
        double* gl = xs + (size_t)(wave + 1) * d_pad;
        if (in_regs) {
#pragma unroll
            for (int j = 0; j < kGradRegs; ++j) {
                const int i = lane + 64 * j;
                if (i < d_pad) {
                    double v = g[0] * av[0][j];
#pragma unroll
                    for (int q = 1; q < kRows; ++q) v = fma(g[q], av[q][j], v);
                    gl[i] = v;
                }
            }
        } else {
            for (int i = lane; i < d_pad; i += 64) {
                double v = g[0] * row0[i];
#pragma unroll
                for (int q = 1; q < kRows; ++q) v = fma(g[q], row0[(size_t)q * d_pad + i], v);
                gl[i] = v;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kRB / 16) {
        // per-16-row partial of sum_s -log sigma_s (samples in order)
        const int h = threadIdx.x;
        double sacc = 0.0;
        for (int q = 0; q < 16; ++q) sacc = sacc + nl[16 * h + q];
        nlog_part[(size_t)c * (m_pad / 16) + bx * (kRB / 16) + h] = sacc;
    }
    if constexpr (kGrad) {
        const double* gl = xs + d_pad;
        gmem<double>* __restrict__ gp = G(PD.gpart) + ((size_t)c * gxb + bx) * d_pad;
        for (int i = threadIdx.x; i < d_pad; i += blockDim.x)
            gp[i] = (gl[i] + gl[d_pad + i]) + (gl[2 * d_pad + i] + gl[3 * d_pad + i]);
    }
}

// ---------------------------------------------------------------------------
// Epilogue shared by the Hessian kernels: each thread holds sixteen entries
// H(i, j) of its wave's 32 x 32 sub-tile (hval, with row iv and column jv).
// Upper entries (i <= j < d) get +lambda on the diagonal; at creation they
// become the learned Hessian, in a round they give Delta = H - learned and the
// tile's partial sums of Delta^2 (off-diagonal, diagonal) for Lk
// (frobeniusNormForSymmetricMatrixFromUpPart, MatrixNMD_SIMD_double.h:1067-1160).
// With with_grad, the diagonal tile also writes the client gradient
// lambda x + (per-column sums gathered in gacc).
// ---------------------------------------------------------------------------
template <bool kInit>
__device__ __forceinline__ void hess_epilogue(const double (&hval)[16], const int (&iv)[16], const int (&jv)[16],
                                              int c, int d, int d_pad, double lambda, size_t S, int tiles_up, int t,
                                              const gmem<double>* __restrict__ x, gmem<double>* __restrict__ learned,
                                              gmem<double>* __restrict__ delta, gmem<double>* __restrict__ hout,
                                              gmem<double>* __restrict__ part,
                                              gmem<double>* __restrict__ grad_c, bool with_grad, double gacc, int i0,
                                              double (*red)[4], double (*gred)[kTile])
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // all sixteen learned values are loaded before use (clamped addresses for
    // entries outside the upper triangle)
    double soff = 0.0, sdiag = 0.0;
    double lv[16];
    if (!kInit) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = iv[q], j = jv[q];
            const bool ok = (j < d && i <= j);
            const size_t o = (size_t)c * S + (ok ? (size_t)j * (j + 1) / 2 + i : 0);
            lv[q] = learned[o];
        }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = iv[q], j = jv[q];
        if (j < d && i <= j) {
            double h = hval[q];
            if (i == j)
                h = h + lambda;
            const size_t o = (size_t)c * S + (size_t)j * (j + 1) / 2 + i;
            if (kInit) {
                learned[o] = h;
            } else {
                const double dv = h - lv[q];
                delta[o] = dv;
                if (hout)
                    hout[o] = h;
                if (i == j)
                    sdiag = fma(dv, dv, sdiag);
                else
                    soff = fma(dv, dv, soff);
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
    if (with_grad)
        gred[tid >> 6][tid & 63] = gacc;
    __syncthreads();
    if (tid == 0) {
        const double o = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        const double dd = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        part[((size_t)c * tiles_up + t) * 2 + 0] = o;
        part[((size_t)c * tiles_up + t) * 2 + 1] = dd;
    }
    if (with_grad && tid < kTile) {
        const int i = i0 + tid;
        if (i < d) {
            const double sum = (gred[0][tid] + gred[1][tid]) + (gred[2][tid] + gred[3][tid]);
            grad_c[(size_t)c * d_pad + i] = lambda * x[i] + sum;
        }
    }
}

// block -> (client, upper tile) with all tiles of one client on one XCD
__device__ inline bool hess_tile_of_block(int b, int n, int tiles_up, int& c, int& t, int& i0, int& j0)
{
    const int xcd = b & 7, slot = b >> 3;
    const int q = slot / tiles_up;
    t = slot - q * tiles_up;
    c = xcd + 8 * q;
    if (c >= n)
        return false;
    int bj = 0;
    while ((bj + 1) * (bj + 2) / 2 <= t) ++bj;
    const int bi = t - bj * (bj + 1) / 2;
    i0 = bi * kTile;
    j0 = bj * kTile;
    return true;
}

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// k_hessian_f32: the same contraction with v_mfma_f32_16x16x4f32 on an FP32
// copy of A (fp32 D * A, FP32 accumulation). Only the contraction is FP32:
// the result is widened to FP64 before lambda, Delta and Lk. The gradient is
// FP64: the diagonal tiles add up the margin pass's partials (or, under
// FedNL-PP, k_grad_f64 computes it). Lane layout of the output: lane (l16, kq),
// item r holds D[m = 4 kq + r][n = l16] (pinned by the layout probe test).
// ---------------------------------------------------------------------------
template <bool kInit>
__global__ __launch_bounds__(256) void k_hessian_f32(FEDNL_KARGS, double* __restrict__ learned_out)
{
    FEDNL_BLOCK();
    const gmem<float>* __restrict__ A32 = G(PD.A32);
    const gmem<double>* __restrict__ Dw = G(PD.Dw);
    const gmem<double>* __restrict__ x = G(PD.x);
    const int n = PD.n, m_pad = PD.m_pad, d = PD.d, d_pad = PD.d_pad, tiles_up = PD.tiles_up;
    const double lambda = PD.lambda;
    const size_t S = PD.S;
    gmem<double>* __restrict__ learned = G(learned_out ? learned_out : PD.learned);
    gmem<double>* __restrict__ delta = G(PD.delta);
    gmem<double>* __restrict__ part = G(PD.part);
    int c, t, i0, j0;
    if (!hess_tile_of_block(blk, n, tiles_up, c, t, i0, j0))
        return;
    if (!kInit && pp_idle(PD, c))
        return;
    __shared__ __align__(16) float sI[2][kKC][kTile + kPadLds];
    __shared__ __align__(16) float sJ[2][kKC][kTile + kPadLds];
    __shared__ double red[2][4];
    __shared__ double gred[4][kTile];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave & 1, wj = wave >> 1;
    const int l16 = lane & 15, kq = lane >> 4;

    flt4 acc[2][2];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int it = 0; it < 2; ++it) acc[jt][it] = flt4{0.f, 0.f, 0.f, 0.f};

    const gmem<float>* Ac = A32 + (size_t)c * m_pad * d_pad;
    const gmem<double>* Dc = Dw + (size_t)c * m_pad;
    // rows lr of a chunk, columns lc .. lc + 3 of both column blocks
    const int lr = tid >> 4, lc = (tid & 15) * 4;
    float4 ri, rj;
    float rd;
    auto gload = [&](int s0) {
        const gmem<float>* row = Ac + (size_t)(s0 + lr) * d_pad;
        ri = *reinterpret_cast<const gmem<float4>*>(row + i0 + lc);
        rj = *reinterpret_cast<const gmem<float4>*>(row + j0 + lc);
        rd = (float)Dc[s0 + lr];
    };
    auto lstore = [&](int buf) {
        *reinterpret_cast<float4*>(&sI[buf][lr][lc]) = ri;
        float4 sj;
        sj.x = rd * rj.x;
        sj.y = rd * rj.y;
        sj.z = rd * rj.z;
        sj.w = rd * rj.w;
        *reinterpret_cast<float4*>(&sJ[buf][lr][lc]) = sj;
    };
    const int nch = m_pad / kKC;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int buf = ch & 1;
        const bool more = ch + 1 < nch;
        if (more)
            gload((ch + 1) * kKC);
#pragma unroll
        for (int ks = 0; ks < kKC; ks += 4) {
            const int kk = ks + kq;
            const float a0 = sJ[buf][kk][wj * 32 + l16];
            const float a1 = sJ[buf][kk][wj * 32 + 16 + l16];
            const float b0 = sI[buf][kk][wi * 32 + l16];
            const float b1 = sI[buf][kk][wi * 32 + 16 + l16];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (more)
            lstore(buf ^ 1);
        __syncthreads();
    }
    double hval[16];
    int iv[16], jv[16];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int it = 0; it < 2; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = (jt * 2 + it) * 4 + r;
                hval[q] = (double)acc[jt][it][r];
                iv[q] = i0 + wi * 32 + it * 16 + l16;
                jv[q] = j0 + wj * 32 + jt * 16 + 4 * kq + r;
            }
    // the client gradient from the margin pass's 16-row partials (diagonal
    // tiles; wave w adds blocks w, w + 4, ... in order)
    const bool fin_grad = !kInit && PD.gpart != nullptr && i0 == j0;
    double gacc = 0.0;
    if (fin_grad) {
        const int nb = m_pad / 16;
        const gmem<double>* gp = G(PD.gpart) + (size_t)c * nb * d_pad + i0 + lane;
        for (int b = wave; b < nb; b += 4) gacc = gacc + gp[(size_t)b * d_pad];
    }
    hess_epilogue<kInit>(hval, iv, jv, c, d, d_pad, lambda, S, tiles_up, t, x, learned, delta, G(PD.pp_h), part,
                         G(PD.grad_c), fin_grad, gacc, i0, red, gred);
}

// ---------------------------------------------------------------------------
// k_hessian_bf16: split-precision contraction with v_mfma_f32_16x16x32_bf16.
// A is held transposed as two BF16 arrays (hi = bf16(a), lo = bf16(a - hi));
// each chunk of 32 samples forms fp32 D * A for the column block J and splits
// it the same way. H ~= (DA)_hi A_hi + (DA)_lo A_hi + (DA)_hi A_lo, FP32
// accumulation; the lo * lo term is dropped. Output layout as k_hessian_f32.
// ---------------------------------------------------------------------------
constexpr int kKB = 32;        // samples per chunk (K of one bf16 MFMA)
static_assert(kKB % 16 == 0, "k_margin takes m_pad (a multiple of kKB) in blocks of 16 rows");
constexpr int kKBPad = 40;     // LDS row length in bf16 (80 B rows)

__device__ inline void split_bf16(float v, __bf16& hi, __bf16& lo)
{
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}

template <bool kInit>
__global__ __launch_bounds__(256) void k_hessian_bf16(FEDNL_KARGS, double* __restrict__ learned_out)
{
    FEDNL_BLOCK();
    const gmem<__bf16>* __restrict__ AThi = G(PD.AThi);
    const gmem<__bf16>* __restrict__ ATlo = G(PD.ATlo);
    const gmem<double>* __restrict__ Dw = G(PD.Dw);
    const gmem<double>* __restrict__ x = G(PD.x);
    const int n = PD.n, m_pad = PD.m_pad, d = PD.d, d_pad = PD.d_pad, tiles_up = PD.tiles_up;
    const double lambda = PD.lambda;
    const size_t S = PD.S;
    gmem<double>* __restrict__ learned = G(learned_out ? learned_out : PD.learned);
    gmem<double>* __restrict__ delta = G(PD.delta);
    gmem<double>* __restrict__ part = G(PD.part);
    int c, t, i0, j0;
    if (!hess_tile_of_block(blk, n, tiles_up, c, t, i0, j0))
        return;
    if (!kInit && pp_idle(PD, c))
        return;
    __shared__ __align__(16) __bf16 tIh[2][kTile][kKBPad];
    __shared__ __align__(16) __bf16 tIl[2][kTile][kKBPad];
    __shared__ __align__(16) __bf16 tJh[2][kTile][kKBPad];
    __shared__ __align__(16) __bf16 tJl[2][kTile][kKBPad];
    __shared__ double red[2][4];
    __shared__ double gred[4][kTile];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave & 1, wj = wave >> 1;
    const int l16 = lane & 15, kq = lane >> 4;

    flt4 acc[2][2];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int it = 0; it < 2; ++it) acc[jt][it] = flt4{0.f, 0.f, 0.f, 0.f};

    const size_t cbase = (size_t)c * d_pad * m_pad;
    const gmem<double>* Dc = Dw + (size_t)c * m_pad;
    // column col of each block, samples ks .. ks + 7 of the chunk
    const int col = tid >> 2, ks = (tid & 3) * 8;
    uint4 rih, ril, rjh, rjl;
    double rdv[8];
    auto gload = [&](int s0) {
        rih = *reinterpret_cast<const gmem<uint4>*>(AThi + cbase + (size_t)(i0 + col) * m_pad + s0 + ks);
        ril = *reinterpret_cast<const gmem<uint4>*>(ATlo + cbase + (size_t)(i0 + col) * m_pad + s0 + ks);
        rjh = *reinterpret_cast<const gmem<uint4>*>(AThi + cbase + (size_t)(j0 + col) * m_pad + s0 + ks);
        rjl = *reinterpret_cast<const gmem<uint4>*>(ATlo + cbase + (size_t)(j0 + col) * m_pad + s0 + ks);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            rdv[u] = Dc[s0 + ks + u];
    };
    auto lstore = [&](int buf) {
        *reinterpret_cast<uint4*>(&tIh[buf][col][ks]) = rih;
        *reinterpret_cast<uint4*>(&tIl[buf][col][ks]) = ril;
        const bf16x8 jh = *reinterpret_cast<const bf16x8*>(&rjh);
        const bf16x8 jl = *reinterpret_cast<const bf16x8*>(&rjl);
        bf16x8 oh, ol;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float a = (float)jh[u] + (float)jl[u];
            __bf16 h, l;
            split_bf16((float)rdv[u] * a, h, l);
            oh[u] = h;
            ol[u] = l;
        }
        *reinterpret_cast<bf16x8*>(&tJh[buf][col][ks]) = oh;
        *reinterpret_cast<bf16x8*>(&tJl[buf][col][ks]) = ol;
    };
    const int nch = m_pad / kKB;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int buf = ch & 1;
        const bool more = ch + 1 < nch;
        if (more)
            gload((ch + 1) * kKB);
        bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            ah[u] = *reinterpret_cast<const bf16x8*>(&tJh[buf][wj * 32 + u * 16 + l16][8 * kq]);
            al[u] = *reinterpret_cast<const bf16x8*>(&tJl[buf][wj * 32 + u * 16 + l16][8 * kq]);
            bh[u] = *reinterpret_cast<const bf16x8*>(&tIh[buf][wi * 32 + u * 16 + l16][8 * kq]);
            bl[u] = *reinterpret_cast<const bf16x8*>(&tIl[buf][wi * 32 + u * 16 + l16][8 * kq]);
        }
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int it = 0; it < 2; ++it) {
                acc[jt][it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[jt], bh[it], acc[jt][it], 0, 0, 0);
                acc[jt][it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[jt], bl[it], acc[jt][it], 0, 0, 0);
                acc[jt][it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[jt], bh[it], acc[jt][it], 0, 0, 0);
            }
        if (more)
            lstore(buf ^ 1);
        __syncthreads();
    }
    double hval[16];
    int iv[16], jv[16];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int it = 0; it < 2; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = (jt * 2 + it) * 4 + r;
                hval[q] = (double)acc[jt][it][r];
                iv[q] = i0 + wi * 32 + it * 16 + l16;
                jv[q] = j0 + wj * 32 + jt * 16 + 4 * kq + r;
            }
    // the client gradient from the margin pass's 16-row partials (diagonal
    // tiles; wave w adds blocks w, w + 4, ... in order)
    const bool fin_grad = !kInit && PD.gpart != nullptr && i0 == j0;
    double gacc = 0.0;
    if (fin_grad) {
        const int nb = m_pad / 16;
        const gmem<double>* gp = G(PD.gpart) + (size_t)c * nb * d_pad + i0 + lane;
        for (int b = wave; b < nb; b += 4) gacc = gacc + gp[(size_t)b * d_pad];
    }
    hess_epilogue<kInit>(hval, iv, jv, c, d, d_pad, lambda, S, tiles_up, t, x, learned, delta, G(PD.pp_h), part,
                         G(PD.grad_c), fin_grad, gacc, i0, red, gred);
}

// ---------------------------------------------------------------------------
// k_grad_f64: client gradients lambda x + sum_s A_c[s] g_s in FP64
// (logistic_regression.h:148-161) for the reduced-precision Hessian modes
// under FedNL-PP (every client's gradient, also for clients that sit the
// round out). Block (column group of 64, client); four sample phases per
// column, summed in a fixed order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grad_f64(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    const gmem<double>* __restrict__ A = G(PD.A);
    const gmem<double>* __restrict__ gw = G(PD.gw);
    const gmem<double>* __restrict__ x = G(PD.x);
    const int m_pad = PD.m_pad, d = PD.d, d_pad = PD.d_pad;
    const double lambda = PD.lambda;
    gmem<double>* __restrict__ grad_c = G(PD.grad_c);
    __shared__ double gp[4][kTile];
    const int gx = d_pad / kTile;
    const int c = blk / gx, i0 = (blk % gx) * kTile;
    const int ii = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const gmem<double>* Ac = A + (size_t)c * m_pad * d_pad;
    const gmem<double>* Gc = gw + (size_t)c * m_pad;
    double acc = 0.0;
    for (int s0 = ph * 8; s0 < m_pad; s0 += 32) {
        double av[8], gv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            av[u] = Ac[(size_t)(s0 + u) * d_pad + i0 + ii];
            gv[u] = Gc[s0 + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            acc = fma(av[u], gv[u], acc);
    }
    gp[ph][ii] = acc;
    __syncthreads();
    if (threadIdx.x < kTile) {
        const int i = i0 + ii;
        if (i < d)
            grad_c[(size_t)c * d_pad + i] = lambda * x[i] + ((gp[0][ii] + gp[1][ii]) + (gp[2][ii] + gp[3][ii]));
    }
}

// FP32 copy of A (same layout)
__global__ void k_convert_a(const double* __restrict__ A, size_t total, float* __restrict__ A32)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
        A32[e] = (float)A[e];
}

// BF16 hi / lo of fp32(A), transposed (per client d_pad rows of m_pad
// samples), through LDS one 32-sample x 64-column tile per block so that the
// reads (512 B rows) and the writes (16 B per thread, 64 B per row) coalesce.
// Grid: n * (m_pad / 32) * (d_pad / 64) blocks of 256 threads.
__global__ __launch_bounds__(256) void k_convert_bf16t(const double* __restrict__ A, int m_pad, int d_pad,
                                                       __bf16* __restrict__ AThi, __bf16* __restrict__ ATlo)
{
    __shared__ float tile[32][65];
    const int ti = d_pad / 64, ts = m_pad / 32;
    const int c = blockIdx.x / (ti * ts);
    const int r = blockIdx.x - c * ti * ts;
    const int s0 = (r / ti) * 32, i0 = (r % ti) * 64;
    const double* Ac = A + (size_t)c * m_pad * d_pad;
    const int tid = threadIdx.x;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int sr = (tid >> 6) + 4 * u;
        tile[sr][tid & 63] = (float)Ac[(size_t)(s0 + sr) * d_pad + i0 + (tid & 63)];
    }
    __syncthreads();
    const int i = tid >> 2, sq = (tid & 3) * 8;
    bf16x8 h, l;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        __bf16 hh, ll;
        split_bf16(tile[sq + u][i], hh, ll);
        h[u] = hh;
        l[u] = ll;
    }
    const size_t o = (size_t)c * d_pad * m_pad + (size_t)(i0 + i) * m_pad + s0 + sq;
    *reinterpret_cast<bf16x8*>(AThi + o) = h;
    *reinterpret_cast<bf16x8*>(ATlo + o) = l;
}

// ---------------------------------------------------------------------------
// k_hessian_f64: H_c = A_c^T diag(D) A_c on the upper tiles, FP64 MFMA
// block = 4 waves (2 x 2), each wave a 32 x 32 sub-tile of 2 x 2 MFMA 16x16 blocks.
// Blocks are mapped so that all tiles of one client run on one XCD
// (blockIdx % 8 selects the XCD), keeping the client's rows of A in that L2.
// kCh: samples staged per LDS chunk (the tile's K extent), 16 (four
// workgroups per CU) or 32 (two, half the barriers); the engine times both at
// creation. Both issue the same MFMAs in the same order: the same bits.
// ---------------------------------------------------------------------------
template <bool kInit, int kCh>
__global__ __launch_bounds__(256, kCh == 16 ? 4 : 2) void k_hessian_f64(FEDNL_KARGS, double* __restrict__ learned_out)
{
    static_assert(kCh == 16 || kCh == 32, "16 or 32 samples per chunk (m_pad is a multiple of 32)");
    constexpr int kRG = kCh / 16;
    FEDNL_BLOCK();
    const gmem<double>* __restrict__ A = G(PD.A);
    const gmem<double>* __restrict__ Dw = G(PD.Dw);
    const gmem<double>* __restrict__ gw = G(PD.gw);
    const gmem<double>* __restrict__ x = G(PD.x);
    const int n = PD.n, m_pad = PD.m_pad, d = PD.d, d_pad = PD.d_pad, tiles_up = PD.tiles_up;
    const double lambda = PD.lambda;
    const size_t S = PD.S;
    gmem<double>* __restrict__ learned = G(learned_out ? learned_out : PD.learned);
    gmem<double>* __restrict__ delta = G(PD.delta);
    gmem<double>* __restrict__ grad_c = G(PD.grad_c);
    gmem<double>* __restrict__ part = G(PD.part);
    int c, t, i0, j0;
    if (!hess_tile_of_block(blk, n, tiles_up, c, t, i0, j0))
        return;
    if (!kInit && pp_idle(PD, c))
        return;
    const bool dg = (i0 == j0);

    // two LDS buffers: the next chunk's rows are loaded into registers while
    // the current chunk is multiplied, then stored (scaled by D) into the
    // other buffer; one barrier per chunk
    __shared__ __align__(16) double sI[2][kCh][kTile + kPadLds];
    __shared__ __align__(16) double sJ[2][kCh][kTile + kPadLds];
    __shared__ double sG[2][kCh];
    __shared__ double red[2][4];
    __shared__ double gred[4][kTile];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave & 1, wj = wave >> 1;
    const int l16 = lane & 15, kq = lane >> 4;
    const bool below = dg && wi == 1 && wj == 0;

    dbl4 acc[2][2];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int it = 0; it < 2; ++it) acc[jt][it] = dbl4{0.0, 0.0, 0.0, 0.0};
    double gacc = 0.0;

    const gmem<double>* Ac = A + (size_t)c * m_pad * d_pad;
    const gmem<double>* Dc = Dw + (size_t)c * m_pad;
    const gmem<double>* Gc = gw + (size_t)c * m_pad;

    // load assignment: rows lr and lr + 8 of each 16-row group, columns lc,
    // lc + 1 of both column blocks (diagonal tiles read the same block
    // twice, from cache)
    const int lr = tid >> 5, lc = (tid & 31) * 2;
    // one chunk's loads in registers: for each 16-row group q, rows lr and
    // lr + 8, columns lc, lc + 1 of both column blocks, their D and g weights
    double2 ri0[kRG], ri1[kRG], rj0[kRG], rj1[kRG];
    double rd0[kRG], rd1[kRG], rg0[kRG], rg1[kRG];
#define HESS_GLOAD_Q(q, s0_)                                                             \
    {                                                                                    \
        const int r_ = (s0_) + 16 * (q) + lr;                                            \
        const gmem<double>* row0_ = Ac + (size_t)r_ * d_pad;                             \
        const gmem<double>* row1_ = Ac + (size_t)(r_ + 8) * d_pad;                       \
        ri0[q] = *reinterpret_cast<const gmem<double2>*>(row0_ + i0 + lc);               \
        rj0[q] = *reinterpret_cast<const gmem<double2>*>(row0_ + j0 + lc);               \
        ri1[q] = *reinterpret_cast<const gmem<double2>*>(row1_ + i0 + lc);               \
        rj1[q] = *reinterpret_cast<const gmem<double2>*>(row1_ + j0 + lc);               \
        rd0[q] = Dc[r_];                                                                 \
        rd1[q] = Dc[r_ + 8];                                                             \
        rg0[q] = Gc[r_];                                                                 \
        rg1[q] = Gc[r_ + 8];                                                             \
    }
    // the chunk into LDS buffer b_, the J block scaled by D
#define HESS_LSTORE_Q(q, b_)                                                             \
    {                                                                                    \
        const int r_ = 16 * (q) + lr;                                                    \
        *reinterpret_cast<double2*>(&sI[b_][r_][lc]) = ri0[q];                           \
        *reinterpret_cast<double2*>(&sI[b_][r_ + 8][lc]) = ri1[q];                       \
        double2 sj0_, sj1_;                                                              \
        sj0_.x = rd0[q] * rj0[q].x;                                                      \
        sj0_.y = rd0[q] * rj0[q].y;                                                      \
        sj1_.x = rd1[q] * rj1[q].x;                                                      \
        sj1_.y = rd1[q] * rj1[q].y;                                                      \
        *reinterpret_cast<double2*>(&sJ[b_][r_][lc]) = sj0_;                             \
        *reinterpret_cast<double2*>(&sJ[b_][r_ + 8][lc]) = sj1_;                         \
        if (lc == 0) {                                                                   \
            sG[b_][r_] = rg0[q];                                                         \
            sG[b_][r_ + 8] = rg1[q];                                                     \
        }                                                                                \
    }
#define HESS_GLOAD(s0_)                                                                  \
    HESS_GLOAD_Q(0, s0_)                                                                 \
    if constexpr (kRG > 1) HESS_GLOAD_Q(kRG - 1, s0_)
#define HESS_LSTORE(b_)                                                                  \
    HESS_LSTORE_Q(0, b_)                                                                 \
    if constexpr (kRG > 1) HESS_LSTORE_Q(kRG - 1, b_)
    const int nch = m_pad / kCh;
    HESS_GLOAD(0)
    HESS_LSTORE(0)
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int buf = ch & 1;
        const bool more = ch + 1 < nch;
        if (more) {
            HESS_GLOAD((ch + 1) * kCh)
        }
        // on a diagonal tile the (wi = 1, wj = 0) quadrant lies below the
        // diagonal: that wave stages and reduces but multiplies nothing
        if (!below)
#pragma unroll
        for (int ks = 0; ks < kCh; ks += 4) {
            const int kk = ks + kq;
            const double a0 = sJ[buf][kk][wj * 32 + l16];
            const double a1 = sJ[buf][kk][wj * 32 + 16 + l16];
            const double b0 = sI[buf][kk][wi * 32 + l16];
            const double b1 = sI[buf][kk][wi * 32 + 16 + l16];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (!kInit && dg) {
            // wave g4 takes samples 4 g4 .. 4 g4 + 3 of each 16-row group
            const int ii = tid & 63, g4 = tid >> 6;
#pragma unroll
            for (int q = 0; q < kRG; ++q)
#pragma unroll
                for (int kk = 16 * q + g4 * 4; kk < 16 * q + g4 * 4 + 4; ++kk)
                    gacc = fma(sI[buf][kk][ii], sG[buf][kk], gacc);
        }
        if (more) {
            HESS_LSTORE(buf ^ 1)
        }
        __syncthreads();
    }

#undef HESS_GLOAD
#undef HESS_LSTORE
#undef HESS_GLOAD_Q
#undef HESS_LSTORE_Q

    // epilogue: lane (l16, kq), item r of v_mfma_f64_16x16x4f64 holds
    // D[m = 4r + kq][n = l16] (pinned by the layout probe test), i.e.
    // rows i = .. + l16 and columns j = .. + kq + 4r of the sub-tile (jt, it)
    double hval[16];
    int iv[16], jv[16];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int it = 0; it < 2; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = (jt * 2 + it) * 4 + r;
                hval[q] = acc[jt][it][r];
                iv[q] = i0 + wi * 32 + it * 16 + l16;
                jv[q] = j0 + wj * 32 + jt * 16 + kq + 4 * r;
            }
    // FedNL-PP takes every client's gradient from k_grad_f64
    hess_epilogue<kInit>(hval, iv, jv, c, d, d_pad, lambda, S, tiles_up, t, x, learned, delta, G(PD.pp_h), part,
                         grad_c, !kInit && dg && !PD.pp, gacc, i0, red, gred);
}

// ---------------------------------------------------------------------------
// k_margin_sparse: the margin pass of a binary design (every feature 0 or 1)
// from each sample's nonzero columns. The sign-folded row is b_s on those
// columns, so each nonzero term of t_s = sum_i A[s,i] x_i
// (logistic_regression.h:98-101) is b_s x_i; added in column order they are
// the bits of the sequential sum over all columns (its zero terms change
// nothing). Then sigma, the sample weights D and g and the -log sigma
// partials per 16 samples as k_margin; kTrial as there. One lane per sample,
// kSpMarginRows samples per workgroup; dynamic LDS: x.
// ---------------------------------------------------------------------------
constexpr int kSpMarginRows = 256;
template <bool kTrial>
__global__ __launch_bounds__(kSpMarginRows) void k_margin_sparse(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    if (kTrial && (!PD.ls || PD.mst[LS_DONE] != 0.0))
        return;
    if (!kTrial && blk == 0 && threadIdx.x == 0 && PD.round_ctr) {
        // the round starts (the first kernel of every round)
        const int r = *PD.round_ctr;
        if (r < PD.trace_cap)
            PD.trace[(size_t)r * kTraceStride + kTraceT0] = (double)wall_clock64();
    }
    const int m = PD.m, m_pad = PD.m_pad, d = PD.d, d_pad = PD.d_pad;
    const double inv_m = 1.0 / (double)m;
    const int wpc = (m_pad + kSpMarginRows - 1) / kSpMarginRows;
    const int c = blk / wpc, bx = blk - c * wpc;
    const int tid = threadIdx.x;
    const int s = bx * kSpMarginRows + tid;
    extern __shared__ double xs[];
    __shared__ double nl[kSpMarginRows];
    // the row's column range and its label sign, loaded with x
    const gmem<uint32_t>* rp = G(PD.sp_rowptr) + (size_t)c * (m + 1);
    uint32_t k0 = 0, k1 = 0;
    double bs = 0.0;
    if (s < m) {
        k0 = rp[s];
        k1 = rp[s + 1];
        bs = G(PD.sp_b)[(size_t)c * m_pad + s];
    }
    {
        const gmem<double>* __restrict__ x = G(PD.x);
        for (int i = tid; i < d_pad; i += kSpMarginRows) xs[i] = (i < d) ? x[i] : 0.0;
    }
    __syncthreads();
    const gmem<uint16_t>* cols = G(PD.sp_cols);
    double acc = 0.0;
    // eight columns at a time: indices, then x, then the ordered adds
    for (uint32_t k = k0; k < k1; k += 8) {
        int ci[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) ci[u] = cols[k + u < k1 ? k + u : k];
        double xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = xs[ci[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k + u < k1)
                acc = acc + bs * xv[u];
    }
    if (s < m) {
        // logistic_regression.h:98-101,148-161,398-404
        const double sg = 1.0 / (1.0 + exp(-acc));
        if (!kTrial) {
            const double z1 = 1.0 - sg;
            const double z2 = z1 * sg;
            G(PD.Dw)[(size_t)c * m_pad + s] = inv_m * z2;
            G(PD.gw)[(size_t)c * m_pad + s] = z1 * (-inv_m);
        }
        nl[tid] = -log(sg);
    } else {
        if (!kTrial && s < m_pad) {
            G(PD.Dw)[(size_t)c * m_pad + s] = 0.0;
            G(PD.gw)[(size_t)c * m_pad + s] = 0.0;
        }
        nl[tid] = 0.0;
    }
    __syncthreads();
    if (tid < kSpMarginRows / 16) {
        // per-16-row partial of sum_s -log sigma_s (samples in order)
        const int h = tid;
        if (bx * kSpMarginRows + 16 * h < m_pad) {
            double sacc = 0.0;
            for (int q = 0; q < 16; ++q) sacc = sacc + nl[16 * h + q];
            G(PD.nlog)[(size_t)c * (m_pad / 16) + bx * (kSpMarginRows / 16) + h] = sacc;
        }
    }
}

// ---------------------------------------------------------------------------
// k_hessian_sparse: the FP64 client Hessians of a binary design (every feature
// 0 or 1). With labels b_s = +-1 the sign-folded entries are 0 or b_s, so every
// nonzero term A[s,i] (D_s A[s,j]) of the sum over samples
// (logistic_regression.h:365-547) is b_s (D_s b_s) = D_s exactly: H_c[i,j] is
// the sum of D_s over the samples that hold both features, added in ascending
// sample order, which is the bits of the sequential sum over all samples (its
// zero terms change nothing). Each entry's samples are listed at creation.
//
// One workgroup per segment of kSpSeg consecutive packed upper entries of a
// client (nparts segments per client, dispatched heaviest first; the blocks
// of a client on one XCD). The segment's token stream holds, lane after lane,
// the items each lane sums (balanced over the lanes at creation):
//   header  bits 0-11 the entry's offset in the segment, bits 12-15 a code:
//           1..12  an off-diagonal entry with that many samples;
//           13     an off-diagonal entry, its sample count in the next token;
//           14     a diagonal entry, its sample count in the next token;
//           15     the client gradient at a column whose diagonal entry lies
//                  in the segment: count, then the column, in the next two;
//   then the item's sample indices, ascending; or
//           0      an item whose samples are consecutive (a dense column's):
//                  its kind (0 off-diagonal, 1 diagonal, 2 gradient), count,
//                  first sample and, for a gradient, the column.
// Off-diagonal entries without a common sample are not listed (their H is 0);
// every diagonal entry is, and so is every column's gradient
// lambda x_i + sum_s b_s g_s over the same samples (logistic_regression.h:
// 148-161, in the order of that sequential sum). The lanes add into LDS; then
// the segment's Delta = H - learned (or, at creation, learned = H), its Lk
// partials and (FedNL-PP option b) H_c are written in list order.
// Dynamic LDS: H (kSpSeg), D, b g (m_pad each), x (d_pad), the lane offsets,
// a diagonal flag per entry, the stream.
// ---------------------------------------------------------------------------

// Workgroup barrier for LDS data only: the wave's LDS operations complete,
// then s_barrier. Unlike __syncthreads(), whose release fence also waits for
// the wave's global stores, it leaves stores issued before it in flight.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

constexpr int kSpSeg = 2048;
constexpr int kSpThreads = 256;             // threads per workgroup, one token lane each
constexpr int kSpLanes = kSpThreads;
constexpr int kSpLaneWords = kSpLanes + 4;  // lane offsets in LDS (8-byte aligned)
__host__ __device__ inline size_t sparse_lds_bytes(int m_pad, int d_pad, size_t max_tokens)
{
    // (the stream gets eight tokens of slack: a lane reads up to four past its last header)
    return sizeof(double) * ((size_t)kSpSeg + 2 * (size_t)m_pad + (size_t)d_pad) + sizeof(uint32_t) * kSpLaneWords +
           (size_t)kSpSeg + sizeof(uint16_t) * ((max_tokens + 3) / 4 * 4 + 8);
}

// acc plus the weights W of the first rem (< kN) samples listed at sp, in
// order: the indices, then the weights, then the adds (past rem masked)
template <int kN>
__device__ __forceinline__ double sparse_sum_masked(double acc, const uint16_t* sp, uint32_t rem, const double* W)
{
    int sx[kN];
#pragma unroll
    for (int u = 0; u < kN; ++u) sx[u] = sp[(uint32_t)u < rem ? u : 0];
    double wv[kN];
#pragma unroll
    for (int u = 0; u < kN; ++u) wv[u] = W[sx[u]];
#pragma unroll
    for (int u = 0; u < kN; ++u)
        if ((uint32_t)u < rem)
            acc = acc + wv[u];
    return acc;
}

template <bool kInit>
__global__ __launch_bounds__(kSpThreads) void k_hessian_sparse(FEDNL_KARGS, double* __restrict__ learned_out)
{
    FEDNL_BLOCK();
    const int n = PD.n, m_pad = PD.m_pad, d_pad = PD.d_pad, nparts = PD.nparts;
    const double lambda = PD.lambda;
    const size_t S = PD.S;
    // block -> (client, dispatch rank), all segments of a client on one XCD;
    // the rank's record names the segment
    const int xcd = blk & 7, slot = blk >> 3;
    const int q8 = slot / nparts;
    const int rank = slot - q8 * nparts;
    const int c = xcd + 8 * q8;
    if (c >= n)
        return;
    if (!kInit && pp_idle(PD, c))
        return;
    constexpr int kWaves = kSpThreads / 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const size_t seg = (size_t)c * nparts + rank;  // record, lane table and Lk partial slot
    // diagnostic build: per-segment stamps (start, staged, summed, end)
    PROBE_STAMP_AT(6000 + 4 * (seg % 2000), 0);
    const uint4 rec = G(PD.sp_seg)[seg];
    const int t = (int)rec.x;
    const size_t e0 = (size_t)t * kSpSeg;
    const int nent = (int)((S - e0) < (size_t)kSpSeg ? (S - e0) : (size_t)kSpSeg);
    const bool with_grad = !kInit && !PD.pp;
    gmem<double>* __restrict__ learned = G(learned_out ? learned_out : PD.learned);
    gmem<double>* __restrict__ delta = G(PD.delta);
    gmem<double>* __restrict__ hout = G(PD.pp_h);
    gmem<double>* __restrict__ part = G(PD.part);
    constexpr int kPer = kSpSeg / kSpThreads;
    const size_t base = (size_t)c * S + e0;
    extern __shared__ __align__(16) unsigned char smem[];
    double* Hs = reinterpret_cast<double*>(smem);
    double* Dl = Hs + kSpSeg;
    double* Gl = Dl + m_pad;
    double* Xl = Gl + m_pad;
    uint32_t* Lo = reinterpret_cast<uint32_t*>(Xl + d_pad);
    unsigned char* Df = reinterpret_cast<unsigned char*>(Lo + kSpLaneWords);  // diagonal entry flags
    uint16_t* St = reinterpret_cast<uint16_t*>(Df + kSpSeg);
    __shared__ double red[2][kWaves];

    // staging: every load first (the weights D and b g, x, the lane offsets,
    // the stream as 8-byte words), then the learned values this thread
    // updates at the end, then the LDS writes, which wait for the staging
    // loads only; the learned loads stay in flight through the sums. Sizes
    // beyond the unrolled counts are staged by the loops after.
    constexpr int kStM = 4, kStX = 2, kStL = 2, kStW = 8;
    const gmem<uint32_t>* lane_g = G(PD.sp_lane) + seg * (kSpLanes + 1);
    const gmem<double>* Dc = G(PD.Dw) + (size_t)c * m_pad;
    const gmem<double>* Gc = G(PD.gw) + (size_t)c * m_pad;
    const gmem<double>* Bc = G(PD.sp_b) + (size_t)c * m_pad;
    const gmem<double>* __restrict__ xg = G(PD.x);
    const uint32_t len = rec.y;
    const uint32_t nwords = (len + 3) / 4;  // the stream starts 8-byte aligned, padded to whole words
    const gmem<uint64_t>* sw =
        reinterpret_cast<const gmem<uint64_t>*>(G(PD.sp_stream) + ((uint64_t)rec.w << 32 | rec.z));
    double rd[kStM], rg[kStM], rb[kStM], rx[kStX];
    uint32_t rl[kStL];
    uint64_t rw[kStW];
#pragma unroll
    for (int u = 0; u < kStM; ++u) {
        const int i = tid + kSpThreads * u;
        const int ii = i < m_pad ? i : 0;
        rd[u] = Dc[ii];
        if (with_grad) {
            rg[u] = Gc[ii];
            rb[u] = Bc[ii];
        }
    }
    if (with_grad) {
#pragma unroll
        for (int u = 0; u < kStX; ++u) {
            const int i = tid + kSpThreads * u;
            rx[u] = xg[i < d_pad ? i : 0];
        }
    }
#pragma unroll
    for (int u = 0; u < kStL; ++u) {
        const int i = tid + kSpThreads * u;
        rl[u] = lane_g[i <= kSpLanes ? i : 0];
    }
#pragma unroll
    for (int u = 0; u < kStW; ++u) {
        const uint32_t w = tid + kSpThreads * u;
        rw[u] = sw[w < nwords ? w : 0];
    }
    double lv[kPer];
    if (!kInit) {
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int idx = tid + kSpThreads * r;
            lv[r] = learned[base + (idx < nent ? idx : 0)];
        }
    }
#pragma unroll
    for (int u = 0; u < kStM; ++u) {
        const int i = tid + kSpThreads * u;
        if (i < m_pad) {
            Dl[i] = rd[u];
            if (with_grad)
                Gl[i] = rg[u] * rb[u];  // b_s g_s, exact
        }
    }
    if (with_grad) {
#pragma unroll
        for (int u = 0; u < kStX; ++u) {
            const int i = tid + kSpThreads * u;
            if (i < d_pad)
                Xl[i] = rx[u];
        }
    }
#pragma unroll
    for (int u = 0; u < kStL; ++u) {
        const int i = tid + kSpThreads * u;
        if (i <= kSpLanes)
            Lo[i] = rl[u];
    }
    uint64_t* Sw = reinterpret_cast<uint64_t*>(St);
#pragma unroll
    for (int u = 0; u < kStW; ++u) {
        const uint32_t w = tid + kSpThreads * u;
        if (w < nwords)
            Sw[w] = rw[u];
    }
    for (int i = tid + kSpThreads * kStM; i < m_pad; i += kSpThreads) {
        Dl[i] = Dc[i];
        if (with_grad)
            Gl[i] = Gc[i] * Bc[i];
    }
    if (with_grad)
        for (int i = tid + kSpThreads * kStX; i < d_pad; i += kSpThreads) Xl[i] = xg[i];
    for (uint32_t w = tid + kSpThreads * kStW; w < nwords; w += kSpThreads) Sw[w] = sw[w];
    for (int i = tid; i < kSpSeg; i += kSpThreads) Hs[i] = 0.0;
    for (int i = tid; i < kSpSeg / 4; i += kSpThreads) reinterpret_cast<uint32_t*>(Df)[i] = 0u;
    lds_barrier();
    PROBE_STAMP_AT(6000 + 4 * (seg % 2000) + 1, 0);

    {
#ifdef FEDNL_PROBE
        const unsigned long long tw0 = wall_clock64();
#endif
        gmem<double>* __restrict__ grad_c = G(PD.grad_c) + (size_t)c * d_pad;
        uint32_t p = Lo[tid];
        const uint32_t pe = Lo[tid + 1];
        while (p < pe) {
            // the header with the next four tokens: an off-diagonal entry of at
            // most four samples needs no further index reads
            uint32_t tk[5];
#pragma unroll
            for (int u = 0; u < 5; ++u) tk[u] = St[p + u];
            const uint32_t off = tk[0] & 0xFFFu, code = tk[0] >> 12;
            if (code - 1u < 4u) {
                double dv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) dv[u] = Dl[(uint32_t)u < code ? tk[1 + u] : tk[1]];
                double acc = 0.0;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if ((uint32_t)u < code)
                        acc = acc + dv[u];
                Hs[off] = acc;
                p += 1 + code;
                continue;
            }
            if (code == 0) {
                // a run: kind (0 off-diagonal, 1 diagonal, 2 gradient), count,
                // first sample (then the column, for a gradient); the weights
                // of consecutive samples, no index reads
                const uint32_t kind = tk[1], rc = tk[2], s0 = tk[3];
                const bool gr = kind == 2;
                p += gr ? 5 : 4;
                if (gr && !with_grad)
                    continue;
                const int col = gr ? (int)tk[4] : -1;
                const double* W = gr ? Gl : Dl;
                double acc = gr ? lambda * Xl[col] : 0.0;
                uint32_t k = 0;
                for (; k + 16 <= rc; k += 16) {
                    double wv[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) wv[u] = W[s0 + k + u];
#pragma unroll
                    for (int u = 0; u < 16; ++u) acc = acc + wv[u];
                }
                if (k < rc) {
                    double wv[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) wv[u] = W[s0 + (k + u < rc ? k + u : k)];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        if (k + u < rc)
                            acc = acc + wv[u];
                }
                if (gr) {
                    grad_c[col] = acc;
                } else {
                    Hs[off] = acc;
                    if (kind == 1)
                        Df[off] = 1;
                }
                continue;
            }
            ++p;
            uint32_t cnt = code;
            int col = -1;
            if (code >= 13) {
                cnt = tk[1];
                ++p;
                if (code == 15) {
                    col = tk[2];
                    ++p;
                }
            }
            if (code == 15 && !with_grad) {
                p += cnt;  // no gradient at creation or under FedNL-PP
                continue;
            }
            // one sum per item: H (codes up to 14) or the gradient (15), in
            // the same accumulator
            const bool gr = code == 15;
            double acc = gr ? lambda * Xl[col] : 0.0;
            const double* W = gr ? Gl : Dl;
            const uint16_t* sp = St + p;
            uint32_t k = 0;
            if (cnt >= 16) {
                // batches of sixteen: the next batch's indices are read with
                // this batch's weights, then the ordered adds
                int sn[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) sn[u] = sp[u];
                for (; k + 16 <= cnt; k += 16) {
                    double wv[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) wv[u] = W[sn[u]];
                    if (k + 32 <= cnt) {
#pragma unroll
                        for (int u = 0; u < 16; ++u) sn[u] = sp[k + 16 + u];
                    }
#pragma unroll
                    for (int u = 0; u < 16; ++u) acc = acc + wv[u];
                }
            }
            // the rest (under sixteen): one batch of 16, 8 or 4, past the end masked
            const uint32_t rem = cnt - k;
            if (rem > 8)
                acc = sparse_sum_masked<16>(acc, sp + k, rem, W);
            else if (rem > 4)
                acc = sparse_sum_masked<8>(acc, sp + k, rem, W);
            else if (rem > 0)
                acc = sparse_sum_masked<4>(acc, sp + k, rem, W);
            p += cnt;
            if (gr) {
                grad_c[col] = acc;
            } else {
                Hs[off] = acc;
                if (code == 14)
                    Df[off] = 1;
            }
        }
#ifdef FEDNL_PROBE
        // diagnostic build: each wave's time in the sums (first 200 segments)
        if (lane == 0 && seg < 200)
            g_probe[14000 + kWaves * seg + wave] = wall_clock64() - tw0;
#endif
    }
    // (the gradient stores stay in flight)
    lds_barrier();
    PROBE_STAMP_AT(6000 + 4 * (seg % 2000) + 2, 0);

    // the segment in list order: Delta (or learned), Lk partials
    double soff = 0.0, sdiag = 0.0;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int idx = tid + kSpThreads * r;
        if (idx < nent) {
            double hv = Hs[idx];
            const bool dgl = Df[idx] != 0;
            if (dgl)
                hv = hv + lambda;
            if (kInit) {
                learned[base + idx] = hv;
            } else {
                const double dv = hv - lv[r];
                delta[base + idx] = dv;
                if (hout)
                    hout[base + idx] = hv;
                if (dgl)
                    sdiag = fma(dv, dv, sdiag);
                else
                    soff = fma(dv, dv, soff);
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
    lds_barrier();  // (the Delta stores stay in flight)
    if (tid == 0) {
        double o = 0.0, dd = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; w += 2) {
            o = o + (red[0][w] + red[0][w + 1]);
            dd = dd + (red[1][w] + red[1][w + 1]);
        }
        part[seg * 2 + 0] = o;
        part[seg * 2 + 1] = dd;
    }
    PROBE_STAMP_AT(6000 + 4 * (seg % 2000) + 3, 0);
}

// ---------------------------------------------------------------------------
// dopt's MinHeap<4>::minHeapify (MinHeap.h:255-290), values and ids in LDS.
// After a swap it keeps comparing against the value that moved up.
// ---------------------------------------------------------------------------
__device__ inline double heapify_x(double* hv, uint32_t* hid, int i, int n, double xval, uint32_t xid)
{
    // sifts element (xval, xid) that sits at slot i; returns the value left at slot i
    double lowv = xval;
    double first_moved = xval;
    bool moved = false;
    for (;;) {
        const int first = 4 * i + 1;
        if (first >= n)
            break;
        // all four children (and ids) loaded unconditionally so the reads issue
        // together (the heap arrays carry 4 slots of padding); absent children
        // are masked to +inf and never win
        double cv[4];
        uint32_t ci[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            cv[q] = hv[first + q];
            ci[q] = hid[first + q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cv[q] = (first + q < n) ? cv[q] : __builtin_inf();
        int lowest = -1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (lowv > cv[q]) {
                lowest = q;
                lowv = cv[q];
            }
        }
        if (lowest < 0)
            break;
        // the chosen child moves up into slot i; comparison continues against its value
        hv[i] = lowv;
        hid[i] = ci[lowest];
        if (!moved) {
            first_moved = lowv;
            moved = true;
        }
        i = first + lowest;
    }
    hv[i] = xval;
    hid[i] = xid;
    return moved ? first_moved : xval;
}

__device__ inline double heapify(double* hv, uint32_t* hid, int i, int n)
{
    return heapify_x(hv, hid, i, n, hv[i], hid[i]);
}

// minHeapify from slot i after an element moved up out of it: comparisons run
// against cmpv (the value that moved up), the sifted element (xval, xid) lands
// in the final slot. Returns nothing; the heap arrays are updated in place.
__device__ inline void heapify_cont(double* hv, uint32_t* hid, int i, int n, double cmpv, double xval,
                                    uint32_t xid)
{
    double lowv = cmpv;
    for (;;) {
        const int first = 4 * i + 1;
        if (first >= n)
            break;
        double cv[4];
        uint32_t ci[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            cv[q] = hv[first + q];
            ci[q] = hid[first + q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cv[q] = (first + q < n) ? cv[q] : __builtin_inf();
        int lowest = -1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (lowv > cv[q]) {
                lowest = q;
                lowv = cv[q];
            }
        }
        if (lowest < 0)
            break;
        // one lane stores (every lane holds the same values; the wave's
        // later LDS reads see the store in order)
        if ((threadIdx.x & 63) == 0) {
            hv[i] = lowv;
            hid[i] = ci[lowest];
        }
        i = first + lowest;
    }
    if ((threadIdx.x & 63) == 0) {
        hv[i] = xval;
        hid[i] = xid;
    }
}

// ---------------------------------------------------------------------------
// The top five levels of the 4-ary heap (slots 0..340) held in the registers
// of one wave: slot 0 in every lane (r0v, r0i), level 1 (slots 1..4) in lanes
// 0..3 of (v1, i1), level 2 (slots 5..20) in lanes 0..15 of (v2, i2), level 3
// (slots 21..84) in lanes 0..63 of (v3, i3), level 4 (slots 85..340) four per
// lane in (v4, i4): lane o holds the children of level-3 slot 21 + o. The
// children of the level-L slot with offset o are the lanes 4o..4o+3 of level
// L+1, one DPP quad. Deeper slots stay in LDS.
//
// Every child group's minimum and the first child holding it are kept beside
// the values (m1/a1 for level 1, g2/b2 and g3/b3 per quad, g4/c4 per lane), so
// one level of a sift is a lane read and a compare. admit() is dopt's
// increaseMinimumInMinHeap + minHeapify (MinHeap.h:78-85, 255-290) with the
// same comparison sequence: the child that moves up is the first of its group
// holding the group minimum, when that minimum is below the running
// comparison value. Most admissions stop at the root or one level below it.
// ---------------------------------------------------------------------------
struct RegHeap {
    double r0v, v1, v2, v3, v4[4];
    uint32_t r0i, i1, i2, i3, i4[4];
    // group minima: m1 = min of level 1 (uniform), a1 its first lane; g2, g3 =
    // minimum of the lane's quad, b2, b3 = ballot of lanes equal to it; g4 =
    // minimum of the lane's four level-4 slots, c4 the first of them
    double m1, g2, g3, g4;
    uint64_t b2, b3;
    int a1, c4;
    int n, lane;

    __device__ void refresh1()
    {
        const double t = quad_min(v1);
        m1 = bcast_lane(t, 0);
        a1 = __builtin_ctzll(__ballot(v1 == m1) & 0xFull);
    }
    __device__ void refresh2()
    {
        g2 = quad_min(v2);
        b2 = __ballot(v2 == g2);
    }
    __device__ void refresh3()
    {
        g3 = quad_min(v3);
        b3 = __ballot(v3 == g3);
    }
    __device__ void refresh4()
    {
        double mloc = v4[0];
        int cloc = 0;
#pragma unroll
        for (int c = 1; c < 4; ++c)
            if (v4[c] < mloc) {
                mloc = v4[c];
                cloc = c;
            }
        g4 = mloc;
        c4 = cloc;
    }
    __device__ void refresh_all()
    {
        refresh4();
        refresh3();
        refresh2();
        refresh1();
    }

    __device__ void load(const double* hv, const uint32_t* hid, int n_, int lane_)
    {
        n = n_;
        lane = lane_;
        const double inf = __builtin_inf();
        r0v = hv[0];
        r0i = hid[0];
        v1 = (lane < 4 && 1 + lane < n) ? hv[1 + lane] : inf;
        i1 = (lane < 4 && 1 + lane < n) ? hid[1 + lane] : 0u;
        v2 = (lane < 16 && 5 + lane < n) ? hv[5 + lane] : inf;
        i2 = (lane < 16 && 5 + lane < n) ? hid[5 + lane] : 0u;
        v3 = (21 + lane < n) ? hv[21 + lane] : inf;
        i3 = (21 + lane < n) ? hid[21 + lane] : 0u;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int sl = 85 + 4 * lane + c;
            v4[c] = (sl < n) ? hv[sl] : inf;
            i4[c] = (sl < n) ? hid[sl] : 0u;
        }
        refresh_all();
    }

    __device__ void store(double* hv, uint32_t* hid) const
    {
        if (lane == 0) {
            hv[0] = r0v;
            hid[0] = r0i;
        }
        if (lane < 4 && 1 + lane < n) {
            hv[1 + lane] = v1;
            hid[1 + lane] = i1;
        }
        if (lane < 16 && 5 + lane < n) {
            hv[5 + lane] = v2;
            hid[5 + lane] = i2;
        }
        if (21 + lane < n) {
            hv[21 + lane] = v3;
            hid[21 + lane] = i3;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int sl = 85 + 4 * lane + c;
            if (sl < n) {
                hv[sl] = v4[c];
                hid[sl] = i4[c];
            }
        }
    }

    // level-4 entry c (uniform) of this lane
    __device__ double v4at(int c) const { return c == 0 ? v4[0] : (c == 1 ? v4[1] : (c == 2 ? v4[2] : v4[3])); }
    __device__ uint32_t i4at(int c) const { return c == 0 ? i4[0] : (c == 1 ? i4[1] : (c == 2 ? i4[2] : i4[3])); }

    // level-4 entry c (uniform) of lane o becomes (v, id)
    __device__ void set4(int o, int c, double v, uint32_t id)
    {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool hit = (lane == o) && (q == c);
            v4[q] = hit ? v : v4[q];
            i4[q] = hit ? id : i4[q];
        }
    }

    // value and id of slot sidx (uniform), wherever it is held
    __device__ void get(const double* hv, const uint32_t* hid, int sidx, double& v, uint32_t& id) const
    {
        if (sidx == 0) {
            v = r0v;
            id = r0i;
        } else if (sidx < 5) {
            v = bcast_lane(v1, sidx - 1);
            id = __builtin_amdgcn_readlane(i1, sidx - 1);
        } else if (sidx < 21) {
            v = bcast_lane(v2, sidx - 5);
            id = __builtin_amdgcn_readlane(i2, sidx - 5);
        } else if (sidx < 85) {
            v = bcast_lane(v3, sidx - 21);
            id = __builtin_amdgcn_readlane(i3, sidx - 21);
        } else if (sidx < 341) {
            const int q = sidx - 85;
            v = bcast_lane(v4at(q & 3), q >> 2);
            id = __builtin_amdgcn_readlane(i4at(q & 3), q >> 2);
        } else {
            v = hv[sidx];
            id = hid[sidx];
        }
    }

    // slot sidx left the heap: a register-held slot becomes +inf so that it
    // never wins as a child (LDS slots are bounded by n)
    __device__ void kill(int sidx)
    {
        const double inf = __builtin_inf();
        if (sidx >= 1 && sidx < 5) {
            v1 = (lane == sidx - 1) ? inf : v1;
            refresh1();
        } else if (sidx >= 5 && sidx < 21) {
            v2 = (lane == sidx - 5) ? inf : v2;
            refresh2();
        } else if (sidx >= 21 && sidx < 85) {
            v3 = (lane == sidx - 21) ? inf : v3;
            refresh3();
        } else if (sidx >= 85 && sidx < 341) {
            set4((sidx - 85) >> 2, (sidx - 85) & 3, inf, 0u);
            refresh4();
        }
    }

    // slot 0 receives (key, id) and sifts down
    __device__ void admit(double* hv, uint32_t* hid, double key, uint32_t id)
    {
        if (key > m1)
            admit_deep(hv, hid, key, id);
        else {
            r0v = key;
            r0i = id;
        }
    }

    // admit() for a key above the level-1 minimum: the root takes that
    // minimum and the sift continues below it
    __device__ void admit_deep(double* hv, uint32_t* hid, double key, uint32_t id)
    {
        // level 0 -> 1
        const int o1 = a1;
        r0v = m1;
        r0i = __builtin_amdgcn_readlane(i1, o1);
        double lowv = m1;
        // level 1 -> 2
        double mu = bcast_lane(g2, 4 * o1);
        if (!(lowv > mu)) {
            v1 = (lane == o1) ? key : v1;
            i1 = (lane == o1) ? id : i1;
            refresh1();
            return;
        }
        const int o2 = 4 * o1 + __builtin_ctzll((b2 >> (4 * o1)) & 0xFull);
        {
            const uint32_t cid = __builtin_amdgcn_readlane(i2, o2);
            v1 = (lane == o1) ? mu : v1;
            i1 = (lane == o1) ? cid : i1;
        }
        lowv = mu;
        // level 2 -> 3
        mu = bcast_lane(g3, 4 * o2);
        if (!(lowv > mu)) {
            v2 = (lane == o2) ? key : v2;
            i2 = (lane == o2) ? id : i2;
            refresh2();
            refresh1();
            return;
        }
        const int o3 = 4 * o2 + __builtin_ctzll((b3 >> (4 * o2)) & 0xFull);
        {
            const uint32_t cid = __builtin_amdgcn_readlane(i3, o3);
            v2 = (lane == o2) ? mu : v2;
            i2 = (lane == o2) ? cid : i2;
        }
        lowv = mu;
        // level 3 -> 4: the children of level-3 slot o3 are lane o3's four
        // level-4 entries
        mu = bcast_lane(g4, o3);
        if (!(lowv > mu)) {
            v3 = (lane == o3) ? key : v3;
            i3 = (lane == o3) ? id : i3;
            refresh3();
            refresh2();
            refresh1();
            return;
        }
        const int c4s = __builtin_amdgcn_readlane(c4, o3);
        {
            const uint32_t cid = __builtin_amdgcn_readlane(i4at(c4s), o3);
            v3 = (lane == o3) ? mu : v3;
            i3 = (lane == o3) ? cid : i3;
        }
        lowv = mu;
        // level 4 -> 5 and below, in LDS (all lanes run the same uniform steps)
        const int i = 85 + 4 * o3 + c4s;
        const int first = 4 * i + 1;
        double cv[4];
        uint32_t ci[4];
        if (first < n) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                cv[c] = hv[first + c];
                ci[c] = hid[first + c];
            }
        }
        int lowest = -1;
        if (first < n) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double x = (first + c < n) ? cv[c] : __builtin_inf();
                if (lowv > x) {
                    lowest = c;
                    lowv = x;
                }
            }
        }
        if (lowest < 0) {
            set4(o3, c4s, key, id);
        } else {
            uint32_t cid = ci[0];
#pragma unroll
            for (int c = 1; c < 4; ++c)
                cid = (lowest == c) ? ci[c] : cid;
            set4(o3, c4s, lowv, cid);
            __builtin_amdgcn_wave_barrier();
            heapify_cont(hv, hid, first + lowest, n, lowv, key, id);
            __builtin_amdgcn_wave_barrier();
        }
        refresh4();
        refresh3();
        refresh2();
        refresh1();
    }
};

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

// |v| of `len` consecutive doubles into LDS by threads [t0, t0 + nt), eight
// loads per thread in flight before the stores
__device__ inline void stage_abs(double* dst, const double* __restrict__ src, int len, int t0, int nt)
{
    for (int e0 = t0; e0 < len; e0 += nt * 8) {
        // clamped, unconditional loads: a guarded load would be issued and
        // waited for one at a time
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = e0 + q * nt;
            v[q] = src[e < len ? e : len - 1];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = e0 + q * nt;
            if (e < len)
                dst[e] = fabs(v[q]);
        }
    }
}

// exclusive prefix sum of one value per thread (256 threads)
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off)
            incl += o;
    }
    if (lane == 63)
        wsum[wave] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
        if (w < wave)
            base += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    if (total)
        *total = tot;
    return base + incl - v;
}

// writes the set bits of bm (positions 32*w + b) in ascending order
__device__ inline void compact_bitmap(const uint32_t* bm, size_t words, uint32_t* out)
{
    const size_t per = (words + blockDim.x - 1) / blockDim.x;
    const size_t w0 = (size_t)threadIdx.x * per;
    const size_t w1 = min(w0 + per, words);
    uint32_t local = 0;
    for (size_t w = w0; w < w1; ++w)
        local += __popc(bm[w]);
    uint32_t o = block_exclusive_scan(local, nullptr);
    for (size_t w = w0; w < w1; ++w) {
        uint32_t bits = bm[w];
        while (bits) {
            const int bpos = __builtin_ctz(bits);
            out[o++] = (uint32_t)(w * 32 + bpos);
            bits &= bits - 1;
        }
    }
}

// Lk (TopK/TopLEK worker: frobeniusNormForSymmetricMatrixFromUpPart of Delta,
// LocalWorkerForFedNL_TopK.h:74-78) from the Hessian kernel's tile partials,
// summed in tile order.
__device__ inline double fro2_from_parts(const double* part, int c, int nparts)
{
    // (off-diagonal, diagonal) pairs, eight slots' loads in flight; summed in
    // slot order
    const gmem<double2>* pc = G(reinterpret_cast<const double2*>(part) + (size_t)c * nparts);
    double o = 0.0, dd = 0.0;
    for (int t0 = 0; t0 < nparts; t0 += 8) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            v[u] = pc[t0 + u < nparts ? t0 + u : 0];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (t0 + u < nparts) {
                o = o + v[u].x;
                dd = dd + v[u].y;
            }
    }
    return (o + o) + dd;
}

// ---------------------------------------------------------------------------
// k_select_heap: dopt's getTopKFromUpperDiagonalPart<true> heap scan
// (SpecialMathRoutinesForMatrix.h:276-327) and, for TopLEK, the energy walk of
// getTopLEKFromUpperDiagonalPart<true> (:377-511) up to the random draw.
// One workgroup per client: all four waves build the heap and order the
// selection; wave 0 alone runs the sequential scan, reading |Delta| from
// global memory through a small LDS ring: each chunk of 64 list positions is
// tested against the current root with one ballot, and only admitted
// positions enter the (sequential) heap update.
// Output: sorted selected list positions, count; TopLEK: hit flag, p, popped.
// kWide: the heap lives in the client's global scratch (k above what LDS
// holds); the ring of staged groups stays in LDS and the tail sorts the ids
// in LDS.
// ---------------------------------------------------------------------------
template <bool kLek, bool kWide>
__device__ __forceinline__ void select_heap_client(const ProblemDev& PD, int blk)
{
    const gmem<double>* __restrict__ delta = G(PD.delta);
    const size_t S = PD.S;
    const int k = PD.k, tiles_up = PD.tiles_up;
    const gmem<double>* __restrict__ part = G(PD.part);
    const double lek_delta = PD.lek_delta;
    gmem<uint32_t>* __restrict__ sel = G(PD.sel);
    gmem<int>* __restrict__ sel_cnt = G(PD.sel_cnt);
    gmem<double>* __restrict__ lek_p = G(PD.lek_p);
    gmem<int>* __restrict__ lek_hit = G(PD.lek_hit);
    gmem<uint32_t>* __restrict__ lek_popped = G(PD.lek_popped);
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = blk;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (pp_idle(PD, c)) {
        if (kLek && tid == 0)
            lek_hit[c] = 0;  // no draw from the worker's stream
        return;
    }
    const double* dc = delta + (size_t)c * S;
    __shared__ double s_fro2;
    PROBE_T(t_sel0);
    // TopLEK's energy target needs ||Delta||_F^2 (from the Hessian kernel's
    // tile partials)
    if (kLek && tid == 0)
        s_fro2 = part ? fro2_from_parts(part, c, PD.nparts) : 0.0;
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
    const size_t words = (S + 31) / 32;
    // the tail orders the selection through a bitmap in the value area, or
    // (when that is too small) by a bitonic sort of P ids
    const bool bm_tail = !kWide && select_bitmap_fits(S, k);
    double* hv;      // k + 4 (padding read by heapify)
    uint32_t* hid;   // k + 4, or P + 4 for the sort
    double* ring;    // sel_ring * kScanGroup * 64
    if (kWide) {
        hv = (double*)(PD.sel_scr + (size_t)c * PD.sel_scr_stride);
        hid = (uint32_t*)(hv + k + 4);
        ring = (double*)smem;
    } else {
        hv = (double*)smem;
        hid = (uint32_t*)(hv + k + 4);
        ring = (double*)(hid + round_up((bm_tail ? k : P) + 4, 2));
    }
    __shared__ int s_ready[kMaxScanRing];
    __shared__ int s_cons;
    if (tid < kMaxScanRing)
        s_ready[tid] = 0;
    if (tid == 0)
        s_cons = 0;

    stage_abs(hv, dc, k, tid, blockDim.x);
    for (int q = tid; q < k; q += blockDim.x)
        hid[q] = (uint32_t)q;
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

    PROBE_ADD(1000 + 8 * c + 0, t_sel0);
    PROBE_T(t_sel1);
    // scan of list positions k .. S-1 in groups of kScanGroup chunks of 64
    // positions. Waves 1..3 stream the groups' |Delta| (-1 past the end) into
    // a ring of R slots in LDS, two groups of loads in flight per wave, each
    // slot stamped with its group when written; wave 0 consumes the groups in
    // order: a group none of whose values exceeds the current root admits
    // nothing and is skipped whole, otherwise each chunk is tested with one
    // ballot against the current root and its admissions run in list order,
    // each later position of the chunk tested again against the root left by
    // the admission before it. (The quirky heap's root can also drop after an
    // admission, so no test is ever made against an older root.)
    const int rest = (int)(S - (size_t)k);
    const int ngroups = (rest + kScanGroup * 64 - 1) / (kScanGroup * 64);
    const int R = PD.sel_ring;
    RegHeap rh;
    if (wave == 0) {
        rh.load(hv, hid, k, lane);
#ifdef FEDNL_PROBE
        const unsigned long long tscan = clock64();
        unsigned long long adm_cnt = 0, adm_cyc = 0;
#endif
        // the admissions of one chunk (positions base .. base + 63 in v)
        auto scan_chunk = [&](double v, int base) {
            double rv = rh.r0v;
            uint64_t mask = __ballot(v > rv);
            while (mask) {
                const int f = __builtin_ctzll(mask);
                const double key = bcast_lane(v, f);
                const uint32_t kid = (uint32_t)(k + base + f);
#ifdef FEDNL_PROBE
                const unsigned long long ta = clock64();
#endif
                // increaseMinimumInMinHeap (MinHeap.h:78-85)
                if (key > rh.m1)
                    rh.admit_deep(hv, hid, key, kid);
                else {
                    rh.r0v = key;
                    rh.r0i = kid;
                }
#ifdef FEDNL_PROBE
                adm_cnt += 1;
                adm_cyc += clock64() - ta;
#endif
                rv = rh.r0v;
                const uint64_t above = (f == 63) ? 0ull : (~0ull << (f + 1));
                mask = __ballot(v > rv) & above;
            }
        };
        for (int g = 0; g < ngroups; ++g) {
            const int slot = g % R;
            scan_lds_wait_eq(&s_ready[slot], g + 1);
            const double* cur = ring + (size_t)slot * (kScanGroup * 64);
            double v[kScanGroup];
            double mx = -1.0;
#pragma unroll
            for (int j = 0; j < kScanGroup; ++j) {
                v[j] = cur[64 * j + lane];
                mx = fmax(mx, v[j]);
            }
            // every lane has its values in registers: the slot is free again
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            if (lane == 0)
                *(volatile int*)&s_cons = g + 1;
            if (__ballot(mx > rh.r0v) == 0ull)
                continue;
            const int gbase = g * kScanGroup * 64;
#pragma unroll
            for (int j = 0; j < kScanGroup; ++j) {
                if (__ballot(v[j] > rh.r0v) != 0ull)
                    scan_chunk(v[j], gbase + 64 * j);
            }
        }
#ifdef FEDNL_PROBE
        if (lane == 0) {
            g_probe[1000 + 8 * c + 5] += clock64() - tscan;
            g_probe[1000 + 8 * c + 3] += adm_cnt;
            g_probe[1000 + 8 * c + 4] += adm_cyc;
        }
#endif
    } else {
        // producers: wave w takes groups w-1, w+2, ...; the loads of its next
        // kDepth groups are in flight (in registers, rotated) while it waits
        // for a free slot and writes the current one
        const gdouble* __restrict__ src = (const gdouble*)(dc + k);
        constexpr int kProd = 3;
        constexpr int kDepth = 2;
        auto load_group = [&](int g, double (&va)[kScanGroup]) {
#pragma unroll
            for (int j = 0; j < kScanGroup; ++j) {
                const int e = g * kScanGroup * 64 + 64 * j + lane;
                va[j] = src[e < rest ? e : rest - 1];  // clamped; replaced by -1 below
            }
        };
        double vq[kDepth][kScanGroup];
        const int g0 = wave - 1;
#pragma unroll
        for (int q = 0; q < kDepth; ++q)
            if (g0 + kProd * q < ngroups)
                load_group(g0 + kProd * q, vq[q]);
        for (int gb = g0; gb < ngroups; gb += kProd * kDepth) {
#pragma unroll
            for (int q = 0; q < kDepth; ++q) {
                const int g = gb + kProd * q;
                if (g >= ngroups)
                    break;
                const int slot = g % R;
                // the consumer is done with group g - R (the slot's last tenant)
                scan_lds_wait_ge(&s_cons, g - R + 1);
                double* dst = ring + (size_t)slot * (kScanGroup * 64);
#pragma unroll
                for (int j = 0; j < kScanGroup; ++j) {
                    const int e = g * kScanGroup * 64 + 64 * j + lane;
                    dst[64 * j + lane] = (e < rest) ? fabs(vq[q][j]) : -1.0;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                if (lane == 0)
                    *(volatile int*)&s_ready[slot] = g + 1;
                if (g + kProd * kDepth < ngroups)
                    load_group(g + kProd * kDepth, vq[q]);
            }
        }
    }
    if (wave == 0)
        rh.store(hv, hid);
    __syncthreads();

    PROBE_ADD(1000 + 8 * c + 1, t_sel1);
    PROBE_T(t_sel2);
    int cnt = k;
    if (kLek) {
        // SpecialMathRoutinesForMatrix.h:383-511, in wave 0: the energy terms
        // of the heap (array order, one chunk of 64 at a time, subtracted in
        // order), then extracFromMinHeap (MinHeap.h:125-141) with the top
        // levels still in registers
        __shared__ int s_cnt;
        if (wave == 0) {
            const double fro2 = s_fro2;
            const double one_minus_delta_xf = (1.0 - lek_delta) * fro2;
            const double eps = 1e-1 * fro2;
            // fro2 = 0 (or NaN): every energy term is zero (or the sums are
            // NaN), the acceptance test never holds and the walk extracts the
            // whole heap without a draw; the result, an empty selection, is
            // set at once
            const bool full_walk = !(fro2 > 0.0);
            double numerator = fro2;
            for (int q0 = 0; q0 < (full_walk ? 0 : k); q0 += 64) {
                const int q = q0 + lane;
                double tq = 0.0;
                if (q < k) {
                    const double it = hv[q];
                    tq = pos_is_diag(hid[q]) ? it * it : 2.0 * it * it;
                }
                const int nq = min(64, k - q0);
                for (int j = 0; j < nq; ++j)
                    numerator -= bcast_lane(tq, j);
            }
            int nh = k;
            int hit = 0;
            double pp = 0.0;
            uint32_t popped = 0;
#ifdef FEDNL_PROBE
            const unsigned long long t_walk = clock64();
#endif
            if (full_walk) {
                nh = 0;
                rh.n = 0;
            }
#ifdef FEDNL_PROBE
            unsigned long long pr_get = 0, pr_admit = 0, pr_rest = 0;
#endif
            while (nh != 0) {
#ifdef FEDNL_PROBE
                const unsigned long long pt0 = clock64();
#endif
                const double prev = numerator;
                const uint32_t pid = rh.r0i;
                const double pv = rh.r0v;
                const int last = nh - 1;
                if (last > 0) {
                    // the last element goes to the root and sifts down a heap of last items
                    double lv;
                    uint32_t lid;
                    rh.get(hv, hid, last, lv, lid);
                    rh.kill(last);
                    rh.n = last;
#ifdef FEDNL_PROBE
                    const unsigned long long pt1 = clock64();
                    pr_get += pt1 - pt0;
#endif
                    rh.admit(hv, hid, lv, lid);
#ifdef FEDNL_PROBE
                    pr_admit += clock64() - pt1;
#endif
                } else {
                    rh.n = 0;
                }
                nh = last;
#ifdef FEDNL_PROBE
                const unsigned long long pt2 = clock64();
#endif
                if (pos_is_diag(pid))
                    numerator += pv * pv;
                else
                    numerator += 2.0 * pv * pv;
                const double next = numerator;
                const bool stop = one_minus_delta_xf + eps > prev && one_minus_delta_xf < next + eps;
#ifdef FEDNL_PROBE
                pr_rest += clock64() - pt2;
#endif
                if (stop) {
                    pp = (one_minus_delta_xf - prev + eps) / (next - prev + eps);
                    hit = 1;
                    popped = pid;
                    break;
                }
            }
#ifdef FEDNL_PROBE
            if (lane == 0) {
                g_probe[2000 + 4 * c + 0] += pr_get;
                g_probe[2000 + 4 * c + 1] += pr_admit;
                g_probe[2000 + 4 * c + 2] += pr_rest;
                g_probe[1000 + 8 * c + 6] += (unsigned long long)(k - nh);
                g_probe[1000 + 8 * c + 7] += clock64() - t_walk;
            }
#endif
            rh.store(hv, hid);
            if (lane == 0) {
                s_cnt = nh;
                lek_hit[c] = hit;
                lek_p[c] = pp;
                lek_popped[c] = popped;
            }
        }
        __syncthreads();
        cnt = s_cnt;
    }

    if (bm_tail) {
        // ascending order through a bitmap of the selected list positions, in
        // the heap value area (not read past this point)
        uint32_t* bm = (uint32_t*)hv;
        for (size_t w = tid; w < words; w += blockDim.x)
            bm[w] = 0u;
        __syncthreads();
        for (int q = tid; q < cnt; q += blockDim.x)
            atomicOr(&bm[hid[q] >> 5], 1u << (hid[q] & 31));
        __syncthreads();
        compact_bitmap(bm, words, sel + (size_t)c * k);
    } else {
        // the ids sorted in LDS (wide: copied there from the global heap,
        // over the ring, which is no longer read)
        uint32_t* srt = kWide ? (uint32_t*)smem : hid;
        for (int q = tid; q < P; q += blockDim.x)
            if (kWide || q >= cnt)
                srt[q] = q < cnt ? hid[q] : 0xFFFFFFFFu;
        __syncthreads();
        bitonic_sort_u32(srt, P);
        for (int q = tid; q < cnt; q += blockDim.x)
            sel[(size_t)c * k + q] = srt[q];
    }
    if (tid == 0)
        sel_cnt[c] = cnt;
    PROBE_ADD(1000 + 8 * c + 2, t_sel2);
}

// ---------------------------------------------------------------------------
// k_select_randk: generateRandKItemsInUpperTriangularPart
// (SpecialMathRoutinesForMatrix.h:95-107, Shuffle.h:31-93) without the
// sequential swaps: every draw u_i comes from an LCG jump-ahead, the swap
// target t_i = i + round(u_i (S-1-i)); the value left at slot i is found by
// following the chain of earlier swaps that touched its source slot.
// kWide: the sort keys and chains in the client's global scratch.
// ---------------------------------------------------------------------------
template <bool kWide>
__device__ __forceinline__ void select_randk_client(const ProblemDev& PD, int blk)
{
    const size_t S = PD.S;
    const int k = PD.k;
    gmem<uint32_t>* __restrict__ rng = G(PD.rng);
    gmem<uint32_t>* __restrict__ sel = G(PD.sel);
    gmem<int>* __restrict__ sel_cnt = G(PD.sel_cnt);
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = blk;
    const int tid = threadIdx.x;
    if (pp_idle(PD, c))
        return;
    const int P = pow2_at_least(k);
    uint64_t* keys = kWide ? (uint64_t*)(PD.sel_scr + (size_t)c * PD.sel_scr_stride) : (uint64_t*)smem;  // P
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
            // every link goes to a smaller index: at most k steps
            int r = src;
            for (int hop = 0; hop < k && pprev[r] >= 0; ++hop) r = pprev[r];
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
__device__ __forceinline__ void select_window_client(const ProblemDev& PD, int blk)
{
    const int comp = PD.comp;
    const size_t S = PD.S;
    const int k = PD.k;
    gmem<uint32_t>* __restrict__ rng = G(PD.rng);
    gmem<uint32_t>* __restrict__ sel = G(PD.sel);
    gmem<int>* __restrict__ sel_cnt = G(PD.sel_cnt);
    gmem<uint32_t>* __restrict__ seq_start = G(PD.seq_start);
    const int c = blk;
    if (pp_idle(PD, c))
        return;
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
// the client's message values, positions, flat indices, selection bitmap and
// learned update (LocalWorkerForFedNL_*.h), after its selection; every
// thread of the workgroup takes part; dynamic LDS holds the bitmap (kWide:
// the bitmap is built in sel_bm itself)
template <bool kWide>
__device__ __forceinline__ void emit_client_body(const ProblemDev& PD, int blk)
{
    const int comp = PD.comp;
    const gmem<double>* __restrict__ delta = G(PD.delta);
    const size_t S = PD.S;
    const int k = PD.k, lda = PD.lda, workers = PD.workers, bm_words = PD.bm_words;
    const gmem<uint32_t>* __restrict__ sel = G(PD.sel);
    const gmem<int>* __restrict__ sel_cnt = G(PD.sel_cnt);
    const gmem<double>* __restrict__ lek_p = G(PD.lek_p);
    const gmem<int>* __restrict__ lek_hit = G(PD.lek_hit);
    const gmem<uint32_t>* __restrict__ lek_popped = G(PD.lek_popped);
    const gmem<uint32_t>* __restrict__ lek_state = G(PD.lek_state);
    const double alpha = PD.alpha, prescale = PD.prescale, mult = PD.mult;
    gmem<double>* __restrict__ learned = G(PD.learned);
    gmem<uint32_t>* __restrict__ pos_out = G(PD.pos);
    gmem<uint32_t>* __restrict__ idx_out = G(PD.idx);
    gmem<double>* __restrict__ vals_out = G(PD.vals);
    gmem<int>* __restrict__ cnt_out = G(PD.cnt);
    gmem<uint32_t>* __restrict__ sel_bm = G(PD.sel_bm);
    extern __shared__ __align__(16) uint32_t bm_lds[];  // bm_words
    const int c = blk;
    if (pp_idle(PD, c))
        return;
    uint32_t* const bml = kWide ? (uint32_t*)(sel_bm + (size_t)c * bm_words) : bm_lds;
    for (int w = threadIdx.x; w < bm_words; w += blockDim.x)
        bml[w] = 0u;
    if (kWide)
        __threadfence();  // the cleared words reach memory before any atomicOr
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
    // eight items per thread at a time: their positions, then their Delta and
    // learned values, all loads of a step in flight together
    constexpr int kEmitBatch = 8;
    for (int t0 = threadIdx.x; t0 < cnt; t0 += blockDim.x * kEmitBatch) {
        uint32_t pv[kEmitBatch];
#pragma unroll
        for (int u = 0; u < kEmitBatch; ++u) {
            const int t = t0 + u * blockDim.x;
            const int tt = t < cnt ? t : cnt - 1;
            uint32_t p;
            if (ins < 0 || tt < ins) p = sc[tt];
            else if (tt == ins) p = s_pop;
            else p = sc[tt - 1];
            pv[u] = p;
        }
        double dv[kEmitBatch], lv[kEmitBatch];
#pragma unroll
        for (int u = 0; u < kEmitBatch; ++u) {
            dv[u] = dc[pv[u]];
            lv[u] = lc[pv[u]];
        }
#pragma unroll
        for (int u = 0; u < kEmitBatch; ++u) {
            const int t = t0 + u * blockDim.x;
            if (t >= cnt)
                break;
            const uint32_t p = pv[u];
            const double item = scaled ? dv[u] * mult : (comp == FEDNL_NATURAL ? natural_round(dv[u]) : dv[u]);
            vals_out[(size_t)c * k + t] = item * prescale;
            const double upd = alpha * item;
            lc[p] = lv[u] + upd;
            pos_out[(size_t)c * k + t] = p;
            idx_out[(size_t)c * k + t] = pos_flat(p, lda);
            atomicOr(&bml[p >> 5], 1u << (p & 31));
        }
    }
    for (int t = cnt + threadIdx.x; t < k; t += blockDim.x) {
        pos_out[(size_t)c * k + t] = 0xFFFFFFFFu;
        idx_out[(size_t)c * k + t] = 0xFFFFFFFFu;
        vals_out[(size_t)c * k + t] = 0.0;
    }
    if (!kWide) {
        __syncthreads();
        for (int w = threadIdx.x; w < bm_words; w += blockDim.x)
            sel_bm[(size_t)c * bm_words + w] = bml[w];
    }
}


template <bool kWide>
__global__ __launch_bounds__(256) void k_emit(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    emit_client_body<kWide>(PD, blk);
}

// kEmit: the client's messages and learned update follow in the same
// workgroup (every compressor but TopLEK, whose draw needs the other clients).
// kWide: heap / sort keys in global scratch, bitmap built in global memory.
template <bool kLek, bool kEmit, bool kWide>
__global__ __launch_bounds__(256) void k_select_heap(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    select_heap_client<kLek, kWide>(PD, blk);
    if (kEmit) {
        __syncthreads();
        emit_client_body<kWide>(PD, blk);
    }
}

template <bool kEmit, bool kWide>
__global__ __launch_bounds__(256) void k_select_randk(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    select_randk_client<kWide>(PD, blk);
    if (kEmit) {
        __syncthreads();
        emit_client_body<kWide>(PD, blk);
    }
}

template <bool kEmit, bool kWide>
__global__ __launch_bounds__(256) void k_select_window(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    select_window_client(PD, blk);
    if (kEmit) {
        __syncthreads();
        emit_client_body<kWide>(PD, blk);
    }
}

// ---------------------------------------------------------------------------
// master_h_block: H[idx] += val for clients 0..n-1 in order (TrainLoop.h:1197-1274).
// Each thread owns one list position and walks the clients in order; a
// client contributes where its selection bitmap has the bit set, with the
// value it sent recomputed from Delta by the same operations as k_emit
// (fl(Delta * prescale), or fl(fl(Delta * S/k) * prescale) for RandK/RandSeqK).
// The bitmap words of the workgroup's 256 positions are staged in LDS; the
// Delta loads of 48 clients are issued together.
// ---------------------------------------------------------------------------
constexpr int kMasterPos = 256;

__device__ __forceinline__ void master_h_block(const ProblemDev& PD, int blk)
{
    gmem<double>* __restrict__ H = G(PD.H);
    const size_t S = PD.S;
    const int n = PD.n, bm_words = PD.bm_words, dp = PD.dp;
    const int scaled = (PD.comp == FEDNL_RANDK || PD.comp == FEDNL_RANDSEQK) ? 1 : 0;
    const bool natural = (PD.comp == FEDNL_NATURAL);
    const gmem<double>* __restrict__ delta = G(PD.delta);
    const gmem<uint32_t>* __restrict__ sel_bm = G(PD.sel_bm);
    const double mult = PD.mult, prescale = PD.prescale;
    gmem<double>* __restrict__ W = G(PD.W);
    // the selection words of the block's positions, client-minor (bwT[q][c]:
    // word q of client c) so that one 16-byte LDS read gives four clients
    __shared__ __align__(16) uint32_t bwT[kMasterPos / 32][kMaxClients + 4];  // +4: store banks
    const size_t p0 = (size_t)blk * kMasterPos;
    const int w0 = (int)(p0 / 32);
    const int nw = min(kMasterPos / 32, bm_words - w0);
    const int n4 = (n + 3) & ~3;  // padded with empty words
    // eight loads per thread in flight before the LDS stores
    for (int e0 = threadIdx.x; e0 < n4 * (kMasterPos / 32); e0 += blockDim.x * 8) {
        uint32_t wv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * blockDim.x;
            const int cc = e / (kMasterPos / 32), q = e % (kMasterPos / 32);
            wv[u] = (cc < n && q < nw && !pp_idle(PD, cc)) ? sel_bm[(size_t)cc * bm_words + w0 + q] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e < n4 * (kMasterPos / 32))
                bwT[e % (kMasterPos / 32)][e / (kMasterPos / 32)] = wv[u];
        }
    }
    __syncthreads();
    const size_t p = p0 + threadIdx.x;
    if (p >= S)
        return;
    const int q = threadIdx.x >> 5;
    const int sh = threadIdx.x & 31;
    double h = H[p];
    // per batch: the clients' words (four per LDS read), then all the Delta
    // loads of the clients whose bit is set, then the sums in client order
    constexpr int kBatch = 48;
    for (int c0 = 0; c0 < n; c0 += kBatch) {
        uint32_t wq[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u += 4) {
            const int cc = min(c0 + u, n4 - 4);
            const uint4 w4 = *reinterpret_cast<const uint4*>(&bwT[q][cc]);
            const bool in = c0 + u < n4;
            wq[u] = in ? w4.x : 0u;
            wq[u + 1] = in ? w4.y : 0u;
            wq[u + 2] = in ? w4.z : 0u;
            wq[u + 3] = in ? w4.w : 0u;
        }
        double v[kBatch];
        bool on[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int cc = c0 + u;
            on[u] = (cc < n) && ((wq[u] >> sh) & 1u);
            v[u] = on[u] ? delta[(size_t)cc * S + p] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (on[u]) {
                const double item = scaled ? v[u] * mult : (natural ? natural_round(v[u]) : v[u]);
                h = h + item * prescale;
            }
        }
    }
    H[p] = h;
    if (W) {
        // the Newton workspace: W(j, i) = H(i, j), column-major lower, ld = dp;
        // k_newton_mc adds the shift on the diagonal (TrainLoop.h:1449-1454)
        const uint32_t j = pos_col(p);
        const uint32_t i = (uint32_t)(p - (size_t)j * (j + 1) / 2);
        W[(size_t)i * dp + j] = h;
    }
}

// averaged gradient (TrainLoop.h:1100-1124,1292-1300): clients in order, 64
// coordinates per workgroup, client gradients staged through LDS
constexpr int kGradCols = 64;
constexpr int kGradClients = 64;

// tile[cc][ii] = src[(c0 + cc) * d_pad + i0 + ii] for cn clients (zero past
// d), all of one thread's loads in flight before the LDS stores (a block of
// 256 threads covers the kGradClients x kGradCols tile in 16 steps)
__device__ __forceinline__ void load_grad_tile(double (*tile)[kGradCols + 1], const gmem<const double>* src, int c0,
                                               int cn, int i0, int d, int d_pad)
{
    constexpr int kSteps = kGradClients * kGradCols / 256;
    const int tot = cn * kGradCols;
    double v[kSteps];
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        const int e = threadIdx.x + u * 256;
        const int cc = e / kGradCols, i = i0 + e % kGradCols;
        v[u] = (e < tot && i < d) ? src[(size_t)(c0 + cc) * d_pad + i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        const int e = threadIdx.x + u * 256;
        if (e < tot)
            tile[e / kGradCols][e % kGradCols] = v[u];
    }
}

__device__ __forceinline__ void master_grad_block(const ProblemDev& PD, int blk)
{
    const int n = PD.n, d = PD.d, d_pad = PD.d_pad;
    const gmem<double>* __restrict__ grad_c = G(PD.grad_c);
    const double inv_n = PD.inv_n;
    gmem<double>* __restrict__ g_avg = G(PD.g_avg);
    __shared__ double tile[kGradClients][kGradCols + 1];
    const int i0 = blk * kGradCols;
    const int col = threadIdx.x & (kGradCols - 1);
    double s = 0.0;
    for (int c0 = 0; c0 < n; c0 += kGradClients) {
        const int cn = min(kGradClients, n - c0);
        load_grad_tile(tile, grad_c, c0, cn, i0, d, d_pad);
        __syncthreads();
        if (threadIdx.x < kGradCols)
            for (int cc = 0; cc < cn; ++cc) s = s + tile[cc][col];
        __syncthreads();
    }
    if (!PD.pp) {
        if (threadIdx.x < kGradCols && i0 + col < d)
            g_avg[i0 + col] = s * inv_n;
        return;
    }
    // FedNL-PP (bin_fednl_local_pp/include/TrainLoop.h:1166,1400-1426): s is
    // the full gradient sum (reported, divided by n); the direction estimate
    // is (sum of the selected clients' changes) / n + the previous estimate
    const gmem<double>* __restrict__ gmsg = G(PD.pp_gmsg);
    double sg = 0.0;
    for (int c0 = 0; c0 < n; c0 += kGradClients) {
        const int cn = min(kGradClients, n - c0);
        load_grad_tile(tile, gmsg, c0, cn, i0, d, d_pad);
        __syncthreads();
        if (threadIdx.x < kGradCols)
            for (int cc = 0; cc < cn; ++cc)
                if (!pp_idle(PD, c0 + cc))
                    sg = sg + tile[cc][col];
        __syncthreads();
    }
    if (threadIdx.x < kGradCols && i0 + col < d) {
        const int i = i0 + col;
        PD.pp_tg[i] = s / (double)n;
        g_avg[i] = sg * inv_n + g_avg[i];
    }
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

// f_c = mean_s(-log sigma_s) + lambda/2 ||x||^2 (logistic_regression.h:119-123)
// from client c's per-block partials of sum_s -log sigma_s
__device__ inline double client_f(const double* __restrict__ nlog_part, int c, int nlog_blocks, int m,
                                  double lambda, double xx)
{
    // the client's partials in block order, eight loads in flight
    const gmem<const double>* pc = G(nlog_part + (size_t)c * nlog_blocks);
    double sacc = 0.0;
    for (int q0 = 0; q0 < nlog_blocks; q0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            v[u] = pc[q0 + u < nlog_blocks ? q0 + u : 0];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (q0 + u < nlog_blocks)
                sacc = sacc + v[u];
    }
    return sacc / (double)m + (lambda / 2.0) * xx;
}

// ---------------------------------------------------------------------------
// master_vec_block: Lk and f averages (TrainLoop.h:1278-1333, 1512-1544), the
// Newton shift, norms for the record (TrainLoop.h:1679-1691), TopLEK stream
// advance. One workgroup.
// ---------------------------------------------------------------------------
// v[0] + v[1] + ... in client order over the clients with on[c] (all when
// on is null), eight reads in flight ahead of the additions
__device__ __forceinline__ double ordered_sum(const double* v, const int* on, int n)
{
    double acc = 0.0;
    for (int c0 = 0; c0 < n; c0 += 8) {
        double t[8];
        int f[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int c = c0 + u < n ? c0 + u : n - 1;
            t[u] = v[c];
            f[u] = (c0 + u < n) && (on == nullptr || on[c] != 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (f[u])
                acc = acc + t[u];
    }
    return acc;
}

__device__ __forceinline__ void master_vec_block(const ProblemDev& PD)
{
    const int comp = PD.comp, n = PD.n, d = PD.d, m = PD.m, nlog_blocks = PD.m_pad / 16, tiles_up = PD.tiles_up;
    const gmem<double>* __restrict__ g_avg = G(PD.g_avg);
    const gmem<double>* __restrict__ part = G(PD.part);
    gmem<double>* __restrict__ lk = G(PD.lk);
    const gmem<double>* __restrict__ nlog_part = G(PD.nlog);
    const gmem<double>* __restrict__ x = G(PD.x);
    const double lambda = PD.lambda, inv_n = PD.inv_n;
    const int option_b = PD.option_b, tracking = PD.tracking, workers = PD.workers, trace_cap = PD.trace_cap;
    const gmem<int>* __restrict__ lek_hit = G(PD.lek_hit);
    gmem<uint32_t>* __restrict__ lek_state = G(PD.lek_state);
    gmem<double>* __restrict__ shift_out = G(PD.shift);
    gmem<double>* __restrict__ trace = G(PD.trace);
    const gmem<int>* __restrict__ round_ctr = G(PD.round_ctr);
    __shared__ double scratch[16];
    __shared__ double fc_s[kMaxClients];
    (void)g_avg;
    PROBE_T(t_v0);
    double xx = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) xx = fma(x[i], x[i], xx);
    xx = block_sum(xx, scratch);
#ifdef FEDNL_PROBE
    if (threadIdx.x == 0) atomicMax(&g_probe[3010], wall_clock64() - t_v0);
#endif
    const bool with_f = tracking || PD.ls;  // f_i is sent (TrainLoop.h:620, 1512-1544)
    const bool pp = PD.pp != 0;
    __shared__ int s_parts, s_items;
    if (threadIdx.x == 0) {
        s_parts = 0;
        s_items = 0;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const double lkc = sqrt(fro2_from_parts(part, c, PD.nparts));
        lk[c] = lkc;
        // this round's participants and message items (integer counts)
        if (!pp || PD.pp_sel[c] != 0) {
            atomicAdd(&s_parts, 1);
            atomicAdd(&s_items, comp == FEDNL_TOPLEK ? PD.cnt[c] : PD.k);
        }
        if (with_f)
            fc_s[c] = client_f(nlog_part, c, nlog_blocks, m, lambda, xx);
    }
    __syncthreads();
#ifdef FEDNL_PROBE
    if (threadIdx.x == 0) atomicMax(&g_probe[3011], wall_clock64() - t_v0);
#endif
    if (threadIdx.x == 0) {
        double l = 0.0, fsum = 0.0;
        l = ordered_sum((const double*)lk, nullptr, n);
        l = l * inv_n;
        if (with_f) {
            // FedNL-PP: the selected clients' f_i over n (TrainLoop.h:1648-1676)
            fsum = ordered_sum(fc_s, pp ? PD.pp_sel : nullptr, n);
            fsum = fsum * inv_n;
        }
        if (PD.ls)
            PD.mst[LS_FXK] = fsum;
        if (PD.pp && option_b) {
            // running Lk average (TrainLoop.h:1436-1445)
            const double sl = ordered_sum(PD.pp_lkmsg, PD.pp_sel, n);
            l = sl * inv_n + PD.mst[PP_LI];
            PD.mst[PP_LI] = l;
        }
        *shift_out = option_b ? l : lambda;  // option a: mu(f) = lambda (logistic_regression.h:702-705)
        const int r = *round_ctr;
        if (r < trace_cap) {
            trace[(size_t)r * kTraceStride + 1] = fsum;
            trace[(size_t)r * kTraceStride + 2] = sqrt(xx);
            trace[(size_t)r * kTraceStride + 5] = (double)s_parts;
            trace[(size_t)r * kTraceStride + 6] = (double)s_items;
        }
#ifdef FEDNL_PROBE
        atomicMax(&g_probe[3012], wall_clock64() - t_v0);
#endif
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
// k_master: the master's aggregation in one launch. Blocks [0, nh) do
// H[idx] += val and W = H, blocks [nh, nh + ng) the gradient average, the last
// block the Lk / f averages, the shift and the record (see the three block
// functions above for the reference lines).
// ---------------------------------------------------------------------------
__host__ __device__ inline int master_h_blocks(size_t S) { return (int)((S + kMasterPos - 1) / kMasterPos); }
__host__ __device__ inline int master_grad_blocks(int d) { return (d + kGradCols - 1) / kGradCols; }

__global__ __launch_bounds__(256) void k_master(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    const int nh = master_h_blocks(PD.S), ng = master_grad_blocks(PD.d);
#ifdef FEDNL_PROBE
    const unsigned long long t_m0 = wall_clock64();
#endif
    if (blk < nh)
        master_h_block(PD, blk);
    else if (blk < nh + ng)
        master_grad_block(PD, blk - nh);
    else
        master_vec_block(PD);
#ifdef FEDNL_PROBE
    if (threadIdx.x == 0)
        atomicMax(&g_probe[3000 + (blk < nh ? 0 : (blk < nh + ng ? 1 : 2))], wall_clock64() - t_m0);
#endif
}

// ---------------------------------------------------------------------------
// GD (--algorithm gd; LocalWorkerForGD.h, TrainLoop.h:993-1052): every client
// sends grad f_c(x) (k_grad_f64), the master averages them in client order
// and steps x -= global_lr * g. k_gd_master: blocks [0, ng) the gradient
// average, the last block f and ||x|| for the record. With tracking the
// reference adds each client's f_c and then sizeof(double) to the aggregate
// before dividing by n (TrainLoop.h:1023-1030), so the recorded f(x) is the
// mean f_c + 8: reproduced, in client order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gd_vec_block(const ProblemDev& PD)
{
    const int n = PD.n, d = PD.d, m = PD.m, nlog_blocks = PD.m_pad / 16;
    const gmem<double>* __restrict__ nlog_part = G(PD.nlog);
    const gmem<double>* __restrict__ x = G(PD.x);
    gmem<double>* __restrict__ trace = G(PD.trace);
    __shared__ double scratch[16];
    __shared__ double fc_s[kMaxClients];
    double xx = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) xx = fma(x[i], x[i], xx);
    xx = block_sum(xx, scratch);
    if (PD.tracking)
        for (int c = threadIdx.x; c < n; c += blockDim.x)
            fc_s[c] = client_f(nlog_part, c, nlog_blocks, m, PD.lambda, xx);
    __syncthreads();
    if (threadIdx.x == 0) {
        double fa = 0.0;
        if (PD.tracking) {
            for (int c = 0; c < n; ++c) {
                fa = fa + fc_s[c];
                fa = fa + (double)sizeof(double);
            }
            fa = fa * PD.inv_n;
        }
        const int r = *PD.round_ctr;
        if (r < PD.trace_cap) {
            trace[(size_t)r * kTraceStride + 1] = fa;
            trace[(size_t)r * kTraceStride + 2] = sqrt(xx);
            trace[(size_t)r * kTraceStride + 5] = (double)n;
            trace[(size_t)r * kTraceStride + 6] = 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void k_gd_master(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    if (blk < master_grad_blocks(PD.d))
        master_grad_block(PD, blk);
    else
        gd_vec_block(PD);
}

// the record's ||g|| (TrainLoop.h:1679-1691), x -= global_lr * g
// (subInPlaceVectorWithMultiple, TrainLoop.h:1050-1051), the round's end
__global__ __launch_bounds__(256) void k_gd_step(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    (void)blk;
    const int d = PD.d;
    const gmem<double>* __restrict__ g = G(PD.g_avg);
    gmem<double>* __restrict__ x = G(PD.x);
    gmem<double>* __restrict__ dx = G(PD.dx);
    gmem<double>* __restrict__ trace = G(PD.trace);
    gmem<int>* __restrict__ round_ctr = G(PD.round_ctr);
    __shared__ double scratch[16];
    const double step = PD.gd_step;
    double gg = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        const double gi = g[i];
        gg = fma(gi, gi, gg);
        const double s = step * gi;
        dx[i] = s;
        x[i] = x[i] - s;
    }
    gg = block_sum(gg, scratch);
    if (threadIdx.x == 0) {
        const int r = *round_ctr;
        if (r < PD.trace_cap) {
            trace[(size_t)r * kTraceStride + 0] = sqrt(gg);
            trace[(size_t)r * kTraceStride + 3] = 1.0;
            trace[(size_t)r * kTraceStride + kTraceT1] = (double)wall_clock64();
        }
        *round_ctr = r + 1;
    }
}


// ---------------------------------------------------------------------------
// FedNL-LS (TrainLoop.h:1546-1640), after k_newton_mc has written dx and left x
// at x_k. k_ls_begin keeps x_k, computes d = -(dx, g) and sets x to the first
// trial x_k - 1 * dx. Each trial is then k_margin<true> (the clients' f_i at
// the trial point) and k_ls_check: f_req = mean_c f_c is accepted when
// f_req <= f(x_k) + (c t) d, else t *= gamma and x = x_k - t dx; after
// max_steps trials x stays at the last one. Once accepted, the remaining
// trial launches return at once. One workgroup per problem.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ls_begin(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    (void)blk;
    if (!PD.ls)
        return;
    const int d = PD.d;
    gmem<double>* __restrict__ x = G(PD.x);
    const gmem<double>* __restrict__ dx = G(PD.dx);
    const gmem<double>* __restrict__ g = G(PD.g_avg);
    gmem<double>* __restrict__ st = G(PD.mst);
    double* __restrict__ xk = st + kLsHead;
    const bool move = PD.ls_max > 0;
    __shared__ double scratch[16];
    double dot = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        const double xi = x[i];
        dot = fma(dx[i], g[i], dot);
        xk[i] = xi;
        if (move)
            x[i] = xi - 1.0 * dx[i];
    }
    dot = block_sum(dot, scratch);
    if (threadIdx.x == 0) {
        st[LS_T] = 1.0;
        st[LS_DD] = -dot;
        st[LS_DONE] = move ? 0.0 : 1.0;
        st[LS_TRIALS] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_ls_check(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    (void)blk;
    if (!PD.ls)
        return;
    if (threadIdx.x == 0) {
        // the last trial launch of the round ends it
        const int r = *PD.round_ctr - 1;
        if (r >= 0 && r < PD.trace_cap)
            PD.trace[(size_t)r * kTraceStride + kTraceT1] = (double)wall_clock64();
    }
    gmem<double>* __restrict__ st = G(PD.mst);
    if (st[LS_DONE] != 0.0)
        return;
    const int n = PD.n, d = PD.d, m = PD.m, nlog_blocks = PD.m_pad / 16;
    gmem<double>* __restrict__ x = G(PD.x);
    const gmem<double>* __restrict__ dx = G(PD.dx);
    const double* __restrict__ xk = st + kLsHead;
    __shared__ double scratch[16];
    __shared__ double fc_s[kMaxClients];
    __shared__ double t_next;
    __shared__ int move;
    // f at the trial point, as k_master computes f(x_k)
    double xx = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) xx = fma(x[i], x[i], xx);
    xx = block_sum(xx, scratch);
    for (int c = threadIdx.x; c < n; c += blockDim.x)
        fc_s[c] = client_f(PD.nlog, c, nlog_blocks, m, PD.lambda, xx);
    __syncthreads();
    if (threadIdx.x == 0) {
        double freq = 0.0;
        for (int c = 0; c < n; ++c) freq = freq + fc_s[c];
        freq = freq * PD.inv_n;
        double t = st[LS_T];
        const int trials = (int)st[LS_TRIALS] + 1;
        const bool accept = freq <= st[LS_FXK] + (PD.ls_c * t) * st[LS_DD];
        if (!accept)
            t = t * PD.ls_gamma;
        const bool done = accept || trials >= PD.ls_max;
        st[LS_T] = t;
        st[LS_TRIALS] = (double)trials;
        st[LS_DONE] = done ? 1.0 : 0.0;
        const int r = *PD.round_ctr - 1;  // k_newton_mc has counted the round
        if (r >= 0 && r < PD.trace_cap)
            PD.trace[(size_t)r * kTraceStride + 4] = (double)trials;
        t_next = t;
        move = done ? 0 : 1;
    }
    __syncthreads();
    if (move) {
        const double t = t_next;
        for (int i = threadIdx.x; i < d; i += blockDim.x) x[i] = xk[i] - t * dx[i];
    }
}

// ---------------------------------------------------------------------------
// FedNL-PP (dopt/bin_fednl_local_pp). k_pp_sample draws the round's clients:
// the first tau entries of a partial shuffle of 0..n-1 (Shuffle.h:31-93)
// with the master's RandomGenIntegerLinear (TrainLoop.h:279-287,1155-1157).
// One workgroup per problem; the draws are sequential in thread 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_pp_sample(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    (void)blk;
    if (!PD.pp)
        return;
    const int n = PD.n, tau = PD.pp_tau;
    __shared__ uint32_t ord[kMaxClients];
    __shared__ int flag[kMaxClients];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        ord[c] = (uint32_t)c;
        flag[c] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t st = *PD.pp_rng;
        for (int i = 0; i < tau; ++i) {
            st = kLcgA * st + kLcgC;
            const double u = (double)st * (1.0 / 4294967295.0);
            const uint64_t off = round_nearest(u * (double)(n - 1 - i));
            if (off) {
                const uint32_t t = ord[i];
                ord[i] = ord[i + off];
                ord[i + off] = t;
            }
        }
        *PD.pp_rng = st;
        for (int q = 0; q < tau; ++q) flag[ord[q]] = 1;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < n; c += blockDim.x) PD.pp_sel[c] = flag[c];
}

// FedNL-PP client step after the learned update (LocalWorkerForFedNL_*.h,
// bin_fednl_local_pp): option b sends Lk_new - Lk_old with
// Lk_new = ||learned_new - H_c(x)||_F over the upper part; every selected
// client sends the change of its direction l_c x + learned_new x - grad f_c(x)
// and keeps the new one. kInit: the starting direction (l_c = 0) of every
// client (LocalWorkerForFedNL_TopK.h:20-30). One workgroup per client.
template <bool kInit>
__global__ __launch_bounds__(256) void k_pp_client(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    const int c = blk;
    if (!PD.pp || (!kInit && pp_idle(PD, c)))
        return;
    const int d = PD.d, d_pad = PD.d_pad;
    const size_t S = PD.S;
    const double* __restrict__ L = PD.learned + (size_t)c * S;
    __shared__ double xs[kMaxD];
    __shared__ double scratch[16];
    __shared__ double s_lk;
    for (int i = threadIdx.x; i < d; i += blockDim.x) xs[i] = PD.x[i];
    double lk_c = 0.0;
    if (!kInit) {
        lk_c = PD.pp_lk[c];
        if (PD.option_b) {
            const double* __restrict__ Hc = PD.pp_h + (size_t)c * S;
            double so = 0.0, sd = 0.0;
            for (size_t p = threadIdx.x; p < S; p += blockDim.x) {
                const double r = L[p] - Hc[p];
                if (pos_is_diag(p))
                    sd = fma(r, r, sd);
                else
                    so = fma(r, r, so);
            }
            so = block_sum(so, scratch);
            sd = block_sum(sd, scratch);
            if (threadIdx.x == 0) {
                const double lk_new = sqrt((so + so) + sd);
                PD.pp_lkmsg[c] = lk_new - PD.pp_lk[c];
                PD.pp_lk[c] = lk_new;
                s_lk = lk_new;
            }
            __syncthreads();
            lk_c = s_lk;
        }
    }
    __syncthreads();
    const double* __restrict__ gc = PD.grad_c + (size_t)c * d_pad;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        // row i of the full symmetric learned Hessian times x, columns ascending
        double mv = 0.0;
        const size_t ri = (size_t)i * (i + 1) / 2;
        for (int j = 0; j < i; ++j) mv = mv + L[ri + j] * xs[j];
        for (int j = i; j < d; ++j) mv = mv + L[(size_t)j * (j + 1) / 2 + i] * xs[j];
        const double dir = (xs[i] * lk_c + mv) - gc[i];
        const size_t o = (size_t)c * d_pad + i;
        if (!kInit)
            PD.pp_gmsg[o] = dir - PD.pp_dir[o];
        PD.pp_dir[o] = dir;
    }
}

// FedNL-PP starting direction estimate: the clients' directions averaged in
// client order (TrainLoop.h:1066-1080)
__global__ __launch_bounds__(256) void k_pp_init_avg(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    (void)blk;
    const int n = PD.n, d = PD.d, d_pad = PD.d_pad;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) s = s + PD.pp_dir[(size_t)c * d_pad + i];
        PD.g_avg[i] = s * PD.inv_n;
    }
}

// ---------------------------------------------------------------------------
// newton_fallback_solve: after a failed factorization (a pivot < 0), dopt's master
// prints "SST FACTORIZATION FAILED" and solves (H + shift I) dx = g by
// Gauss-Jordan elimination with partial pivoting (TrainLoop.h:1410-1420,
// 1476-1486; GaussEliminationSolvers.h:27-120) on its full H, whose lower
// triangle is never updated after creation: M(i, j) = H(i, j) (+ shift on the
// diagonal) for i <= j, H0(j, i) below. M is built row-major in W (ld = dp).
// Per column: the first largest |M(i, col)| at or below the diagonal is
// swapped up, the pivot row and b are scaled (b by division, the row by the
// reciprocal), the column is eliminated from every other row. The products
// and sums are rounded separately, as in the reference's loops. A zero
// tolerance never skips a column, so dx = b at the end. Called by every
// thread of the last k_newton_mc workgroup; kept out of line, off the solve's
// hot path.
// ---------------------------------------------------------------------------
__device__ __noinline__ void newton_fallback_solve(const ProblemDev& PD, double* __restrict__ lds)
{
    const int d = PD.d, ld = PD.dp, tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    gmem<double>* __restrict__ a = G(PD.W);
    const gmem<double>* __restrict__ H = G(PD.H);
    const gmem<double>* __restrict__ H0 = G(PD.H0);
    const double shift = *PD.shift;
    // the solve's LDS: b, the eliminated column, the pivot row, reduction slots
    double* bv = lds;
    double* kcol = lds + ld;
    double* prow = lds + 2 * ld;
    double* red_v = lds + 3 * ld;
    int* red_i = (int*)(lds + 3 * ld + 8);
    int* s_br = red_i + 8;
    for (int e = tid; e < d * d; e += nthr) {
        const int i = e / d, j = e - i * d;
        double v;
        if (i <= j) {
            v = H[(size_t)j * (j + 1) / 2 + i];
            if (i == j)
                v = v + shift;
        } else {
            v = H0[(size_t)i * (i + 1) / 2 + j];
        }
        a[(size_t)i * ld + j] = v;
    }
    for (int i = tid; i < d; i += nthr) bv[i] = PD.nb[i];
    __syncthreads();
    for (int col = 0; col < d; ++col) {
        // pivot: first row of the largest |M(i, col)|, i > col (NaN never wins)
        double bvv = -1.0;
        int bi = d;
        for (int i = col + 1 + tid; i < d; i += nthr) {
            const double v = fabs(a[(size_t)i * ld + col]);
            if (v > bvv) {  // ascending i: the first of equal values stays
                bvv = v;
                bi = i;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bvv, off);
            const int oi = __shfl_xor(bi, off);
            if (ov > bvv || (ov == bvv && oi < bi)) {
                bvv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            red_v[wave] = bvv;
            red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double v = red_v[0];
            int ii = red_i[0];
            for (int w = 1; w < nthr / 64; ++w)
                if (red_v[w] > v || (red_v[w] == v && red_i[w] < ii)) {
                    v = red_v[w];
                    ii = red_i[w];
                }
            const double v0 = fabs(a[(size_t)col * ld + col]);
            *s_br = (ii < d && v > v0) ? ii : col;  // v0 NaN: nothing is larger
        }
        __syncthreads();
        const int br = *s_br;
        if (br != col) {
            for (int j = col + tid; j < d; j += nthr) {
                const double t = a[(size_t)col * ld + j];
                a[(size_t)col * ld + j] = a[(size_t)br * ld + j];
                a[(size_t)br * ld + j] = t;
            }
            if (tid == 0) {
                const double t = bv[col];
                bv[col] = bv[br];
                bv[br] = t;
            }
            __syncthreads();
        }
        const double piv = a[(size_t)col * ld + col];
        const double inv = 1.0 / piv;
        for (int j = col + tid; j < d; j += nthr) {
            const double v = __dmul_rn(a[(size_t)col * ld + j], inv);
            prow[j] = v;
            a[(size_t)col * ld + j] = v;
        }
        for (int i = tid; i < d; i += nthr) kcol[i] = a[(size_t)i * ld + col];
        if (tid == 0)
            bv[col] = bv[col] / piv;
        __syncthreads();
        const double bc = bv[col];
        for (int i = tid; i < d; i += nthr)
            if (i != col)
                bv[i] = __dadd_rn(bv[i], -__dmul_rn(bc, kcol[i]));
        // rows other than col, columns col..d-1 (entries left of col are zero
        // in the pivot row)
        const int w = d - col;
        for (int e = tid; e < (d - 1) * w; e += nthr) {
            int i = e / w;
            const int j = col + (e - i * w);
            i += (i >= col) ? 1 : 0;
            double* pa = a + (size_t)i * ld + j;
            *pa = (j == col) ? 0.0 : __dadd_rn(*pa, __dmul_rn(prow[j], -kcol[i]));
        }
        __syncthreads();
    }
    gmem<double>* __restrict__ x = G(PD.x);
    gmem<double>* __restrict__ dx_out = G(PD.dx);
    for (int i = tid; i < d; i += nthr) {
        dx_out[i] = bv[i];
        if (PD.update_x == 1)
            x[i] = x[i] - bv[i];
        else if (PD.update_x == 2)
            x[i] = bv[i];
    }
}


// Row stride (doubles) of k_newton_mc's block-column rows in LDS. The MFMA operand
// reads take rows l16 = 0..15 and columns kq = 0..3 in one instruction; with
// a stride of 2 mod 32 the 32 lanes of each half-wave hit 32 distinct bank
// pairs.
constexpr int kNewtonPS = kNB + 2;

// 1/sqrt(a) for a > 0: v_rsq_f64 and two Newton steps
__device__ __forceinline__ double rsq_nr2(double a)
{
    double y = __builtin_amdgcn_rsq(a);
    double h = a * y;
    y = fma(0.5 * y, fma(-h, y, 1.0), y);
    h = a * y;
    y = fma(0.5 * y, fma(-h, y, 1.0), y);
    return y;
}

// sqrt(a) from y = rsq_nr2(a): one correction of a y (within an ulp or two
// of the rounded root)
__device__ __forceinline__ double root_from_inv(double a, double y)
{
    double r = a * y;
    return fma(fma(-r, r, a), 0.5 * y, r);
}

// One column step of the diagonal-block factorisation (rows in lanes,
// column kCc, columns up to kEnd - 1): every lane takes the same path. The
// scaled column goes to LDS (cb, double-buffered, 64 slots each) for the
// rank-1 update of columns kCc+2 ..; the entry the next pivot needs,
// L[kCc+1][kCc], goes by v_readlane, so the pivot chain never waits on LDS.
// Lanes at or above the pivot compute entries of the block's upper part,
// which nothing reads. The pivot, its reciprocal root and the column's
// forward-substitution entry (all uniform) go to meta[4 kCc ..].
template <int kCc, int kEnd>
__device__ __forceinline__ void diag_column(double (&row)[kNB], double& yr, bool& fail, double& piv, int lane,
                                            double* __restrict__ cb, double* __restrict__ meta)
{
    // a negative pivot fails the block (reported after it); a zero pivot
    // gives non-finite entries, as a division by it would
    fail = fail || (piv < 0.0);
    meta[4 * kCc + 0] = piv;  // every lane writes the same value
    const double inv = rsq_nr2(piv);
    const double scaled = inv * row[kCc];
    double* b = cb + 64 * (kCc & 1);
    b[lane] = scaled;
    if (kCc + 1 < kEnd) {
        row[kCc + 1] = fma(-scaled, bcast_lane(scaled, kCc + 1), row[kCc + 1]);
        piv = bcast_lane(row[kCc + 1], kCc + 1);
    }
    row[kCc] = scaled;
    const double ycc = bcast_lane(yr, kCc) * inv;
    yr = fma(-scaled, ycc, yr);
    meta[4 * kCc + 1] = inv;
    meta[4 * kCc + 2] = ycc;
    // (the wave's LDS operations complete in order, so the reads below see
    // every lane's scaled entry; no scheduling barrier, so that the next
    // column's pivot chain can overlap these updates)
#pragma unroll
    for (int c2 = kCc + 2; c2 < kEnd; ++c2)
        row[c2] = fma(-scaled, b[c2], row[c2]);
}

template <int kCc, int kEnd>
__device__ __forceinline__ void diag_columns(double (&row)[kNB], double& yr, bool& fail, double& piv, int lane,
                                             double* __restrict__ cb, double* __restrict__ meta)
{
    if constexpr (kCc < kEnd) {
        diag_column<kCc, kEnd>(row, yr, fail, piv, lane, cb, meta);
        diag_columns<kCc + 1, kEnd>(row, yr, fail, piv, lane, cb, meta);
    }
}

// ---------------------------------------------------------------------------
// k_newton_mc: the master's Newton solve (TrainLoop.h:1449-1466;
// CholeskyFactorization.h:153-272; ElementarySolvers.h:72-226) over many
// workgroups, for d up to kMaxMcD.
//
// W holds the lower triangle of H (column-major, ld = dp, dp = d rounded up
// to 32, padding set to the identity); the shift joins the diagonal as the
// columns are loaded. Block column J (columns 32J .. 32J+31) is owned by
// workgroups (J, s), each holding rows 32J + 512 s .. in LDS (left-looking):
//   1. for every finished panel K < J: Col -= P_K[rows] P_K[block J]^T on
//      FP64 MFMA (P_K = block column K of L, read from W);
//   2. (J, 0): its diagonal block is factored by one wave with all 64 lanes:
//      rows 0..31 hold the block, rows 32..63 an identity, so the same column
//      operations give L11 and L11^-T together (the right-hand side is
//      carried along: y_J, the forward substitution);
//   3. every row below the diagonal block becomes Col L11^-T on FP64 MFMA;
//      (J, 0) does the 32 rows of block J+1 first and publishes them with
//      L11, L11^-T and y_J (the next diagonal factor needs only these), then
//      the rest; (J, s > 0) wait for L11^-T.
// The workgroup that finishes last (a ticket) runs L^T dx = y by blocks from
// the bottom (dx_J = L11^-T r_J, then r folded over the rows above) and
// x -= dx. A workgroup only ever waits for lower-numbered ones, which are
// dispatched first, so the launch needs no co-residency. Waits are bounded
// (an expired wait sets the error word). Flags carry an epoch (one per
// solve), so nothing is reset between solves.
// A negative pivot fails the factorisation, as dopt's: the last workgroup
// then solves by Gauss-Jordan elimination (newton_fallback_solve).
// ---------------------------------------------------------------------------
constexpr int kSegRows = 512;                 // rows of a block column one workgroup holds
constexpr int kMcThreads = 512;
constexpr int kMaxMcD = 2048;
constexpr long kMcSpinCap = 1L << 22;         // polls of one wait before it gives up
static_assert(kSegRows == (kMcThreads / 64) * 64, "one 64-row chunk per wave per segment");

__host__ __device__ inline int mc_nseg(int dp, int J) { return (dp - kNB * J + kSegRows - 1) / kSegRows; }
__host__ __device__ inline int mc_blocks(int dp)
{
    int t = 0;
    for (int J = 0; J < dp / kNB; ++J) t += mc_nseg(dp, J);
    return t;
}
__host__ __device__ constexpr int mc_sync_words(int dp) { return 4 + 2 * (dp / kNB); }
__host__ __device__ inline size_t mc_lds_bytes(int dp)
{
    const size_t col = sizeof(double) * (size_t)kNewtonPS * (size_t)(dp < kSegRows ? dp : kSegRows);
    const size_t fallback = sizeof(double) * (3 * (size_t)dp + 16);
    return col > fallback ? col : fallback;
}

// What one workgroup hands to another goes through device-coherent
// (sc1) stores and loads: write-through to memory and never served from a
// non-coherent L2 copy. A flag is stored only after the wave's earlier
// stores have completed (s_waitcnt), and data are loaded only after the
// flag has been seen, so no L2 writeback or invalidate is needed.
__device__ __forceinline__ double ld_dev(const gmem<double>* p)
{
    return __hip_atomic_load((gmem<double>*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(gmem<double>* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// every memory operation of the wave issued so far has completed
__device__ __forceinline__ void mc_drain()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Waits until *p >= v. Every lane of the calling wave runs the same loop.
// Returns false (and sets the error word) if the wait expired.
__device__ inline bool mc_wait_ge(int* p, int v, int* err)
{
    long n = 0;
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) {
        __builtin_amdgcn_s_sleep(1);
        if (++n > kMcSpinCap) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
    }
    mc_drain();
    return n <= kMcSpinCap;
}

// Intra-workgroup signal between waves through LDS.
__device__ inline void mc_lds_signal(volatile int* f, int v)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    *f = v;
}
__device__ inline void mc_lds_wait(volatile int* f, int v)
{
    long n = 0;
    while (*f < v && ++n <= kMcSpinCap)
        __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Col rows 16 rt .. 16 rt + 15 (local) for rt = rt0, rt0 + step, .. < rt1, both
// 16-column tiles (only the lower one where lower_only holds for rt == 0),
// minus P_K[r0 + row][k] P_K[brow + col][k] summed over k < 32, where
// P_K[r][k] = W[(32 K + k) dp + r]. The next tile's loads are in flight
// during the current tile's MFMAs.
__device__ inline void mc_update_tiles(const gmem<double>* __restrict__ W, int dp, int K, int r0, int brow,
                                       double* __restrict__ Col, int rt0, int step, int rt1, bool lower_first,
                                       int lane)
{
    if (rt0 >= rt1)
        return;
    const int l16 = lane & 15, kq = lane >> 4;
    const gmem<double>* __restrict__ PK = W + (size_t)(kNB * K + kq) * dp;
    double av[2][8];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int q = 0; q < 8; ++q)
            av[ct][q] = ld_dev(PK + (size_t)(4 * q) * dp + brow + 16 * ct + l16);
    double bv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        bv[q] = ld_dev(PK + (size_t)(4 * q) * dp + r0 + 16 * rt0 + l16);
    for (int rt = rt0; rt < rt1; rt += step) {
        double bn[8];
        const int rn = (rt + step < rt1) ? rt + step : rt;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            bn[q] = ld_dev(PK + (size_t)(4 * q) * dp + r0 + 16 * rn + l16);
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            if (lower_first && rt == 0 && ct == 1)
                continue;  // above the diagonal of the diagonal block
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 8; ++q)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ct][q], bv[q], acc, 0, 0, 0);
            double* cp = Col + (16 * rt + l16) * kNewtonPS + 16 * ct + kq;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                cp[4 * r] = cp[4 * r] - acc[r];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            bv[q] = bn[q];
    }
}

// Col rows 16 rt .. (local) for rt = rt0, rt0 + step, .. < rt1 become
// Col L11^-T (X = L11^-T, upper, row-major in Xt with stride kNB + 1) on FP64
// MFMA: the triangular solve of the rows below a diagonal block. With Wc
// (the block column's first column at the segment's first row, in W) each
// solved tile also goes to W straight from the accumulators.
__device__ inline void mc_solve_tiles(double* __restrict__ Col, const double (*Xt)[kNB + 1], int rt0, int step,
                                      int rt1, int lane, gmem<double>* __restrict__ Wc = nullptr, int dp = 0)
{
    if (rt0 >= rt1)
        return;
    const int l16 = lane & 15, kq = lane >> 4;
    double av[2][8];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int q = 0; q < 8; ++q)
            av[ct][q] = Xt[4 * q + kq][16 * ct + l16];  // A[m = col][k] = X[k][col]
    for (int rt = rt0; rt < rt1; rt += step) {
        double* rp = Col + (16 * rt + l16) * kNewtonPS + kq;
        double bv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            bv[q] = rp[4 * q];  // B[k][n = row]
        __builtin_amdgcn_wave_barrier();  // the tile's rows are read before they are overwritten
        dbl4 acc[2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 8; ++q)
                acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ct][q], bv[q], acc[ct], 0, 0, 0);
        }
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                rp[16 * ct + 4 * r] = acc[ct][r];
        if (Wc) {
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    st_dev(Wc + (size_t)(16 * ct + 4 * r + kq) * dp + 16 * rt + l16, acc[ct][r]);
        }
    }
}

// The diagonal block's factorisation in one wave: lane r < 32 holds row r of
// the block (Dg), lane 32 + q row q of an identity. The same column
// operations (the lanes execute them anyway) turn the block rows into L11
// and the identity rows into L11^-T (the panel rule X L11^T = I); the
// right-hand side ry is carried along (forward substitution: ry becomes y).
// Two 16-column halves; between them rows 16..63 take -R L21^T on FP64 MFMA
// (R: their first 16 columns, L21: the rows 16..31 among them). On return Dg
// holds L11 (lower), Xt L11^-T (upper), invd 1 / L11[i][i], ry y. A negative
// pivot sets *s_fail. cb (128 doubles) and meta (128 doubles) are LDS
// scratch of the calling wave.
__device__ __forceinline__ void mc_diag_factor(double (*Dg)[kNB + 1], double (*Xt)[kNB + 1], double (*Sq)[17],
                                               double* invd, double* ry, double* cb, double* meta, int lane,
                                               int* s_fail)
{
    constexpr int kHalf = kNB / 2;
    double row[kNB];
    const bool blk = lane < kNB;
    const int lr = lane & (kNB - 1);
    double* const src = blk ? &Dg[lr][0] : &Xt[lr][0];  // Xt holds the identity on entry
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        row[c] = src[c];
    double yr = blk ? ry[lr] : 0.0;
    bool fail = false;
    double piv = bcast_lane(row[0], 0);
    diag_columns<0, kHalf>(row, yr, fail, piv, lane, cb, meta);
    // rows 16..63 (lanes 16..63), first half of their columns, to LDS
    if (lane >= kHalf) {
#pragma unroll
        for (int c = 0; c < kHalf; ++c)
            Sq[lane - kHalf][c] = row[c];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {
        // U[i][c] = sum_k S[i][k] L21[c][k] (L21 = S rows 0..15), three
        // independent 16-row tiles, written back once all are formed
        const int l16 = lane & 15, kq = lane >> 4;
        dbl4 acc[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k0 = 0; k0 < kHalf; k0 += 4)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Sq[l16][k0 + kq], Sq[16 * t + l16][k0 + kq], acc[t],
                                                              0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                Sq[16 * t + l16][4 * r + kq] = acc[t][r];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    {
        const int i = lane >= kHalf ? lane - kHalf : 0;
#pragma unroll
        for (int c = 0; c < kHalf; ++c)
            row[kHalf + c] = (lane >= kHalf) ? row[kHalf + c] - Sq[i][c] : row[kHalf + c];
    }
    piv = bcast_lane(row[kHalf], kHalf);
    diag_columns<kHalf, kNB>(row, yr, fail, piv, lane, cb, meta);
    if (__ballot(fail)) {
        if (lane == 0)
            *s_fail = 1;
    }
    // the per-column records into registers first (meta may alias Dg)
    double m_piv = 0.0, m_inv = 0.0, m_y = 0.0;
    if (blk) {
        m_piv = meta[4 * lr + 0];
        m_inv = meta[4 * lr + 1];
        m_y = meta[4 * lr + 2];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        src[c] = row[c];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (blk) {
        // the diagonal from each column's pivot and reciprocal root
        Dg[lr][lr] = root_from_inv(m_piv, m_inv);
        invd[lr] = m_inv;
        ry[lr] = m_y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (J, 0)'s diagonal block (local row tiles 0, 1) and y_J take panel K's
// block-J rows in one load round trip. Those rows are the MFMA operands of
// P_K[J] P_K[J]^T, and the same registers give y_J -= P_K[J] y_K: lane
// (l16, kq) holds P_K[J][16 ct + l16][4 q + kq], so each lane sums its 8
// products and the four kq lanes of a row are added (kq 0+1, 2+3, then the
// pairs).
__device__ inline void mc_update_diag(const gmem<double>* __restrict__ W, const gmem<double>* __restrict__ ny,
                                      int dp, int K, int brow, double* __restrict__ Col, double* __restrict__ ry,
                                      int lane)
{
    const int l16 = lane & 15, kq = lane >> 4;
    const gmem<double>* __restrict__ PK = W + (size_t)(kNB * K + kq) * dp + brow;
    double av[2][8], yk[8];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int q = 0; q < 8; ++q)
            av[ct][q] = ld_dev(PK + (size_t)(4 * q) * dp + 16 * ct + l16);
#pragma unroll
    for (int q = 0; q < 8; ++q)
        yk[q] = ld_dev(ny + kNB * K + 4 * q + kq);
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            if (rt == 0 && ct == 1)
                continue;  // above the diagonal
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 8; ++q)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ct][q], av[rt][q], acc, 0, 0, 0);
            double* cp = Col + (16 * rt + l16) * kNewtonPS + 16 * ct + kq;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                cp[4 * r] = cp[4 * r] - acc[r];
        }
    double py[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            a = fma(av[ct][q], yk[q], a);
        a += __shfl_xor(a, 16);
        a += __shfl_xor(a, 32);
        py[ct] = a;
    }
    if (kq == 0) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
            ry[16 * ct + l16] = ry[16 * ct + l16] - py[ct];
    }
}

__global__ __launch_bounds__(kMcThreads) void k_newton_mc(FEDNL_KARGS)
{
    FEDNL_BLOCK();
    const int d = PD.d, dp = PD.dp, nb = dp / kNB;
    int J = 0, s = blk;
    while (J < nb - 1 && s >= mc_nseg(dp, J)) {
        s -= mc_nseg(dp, J);
        ++J;
    }
    const int r0 = kNB * J + kSegRows * s;       // global row of Col row 0
    const int nrows = min(kSegRows, dp - r0);    // multiple of 32
    const int ntile = nrows / 16;
    int* const sync = PD.nsync;
    int* const cnt = sync + 4;
    int* const dflag = sync + 4 + nb;
    int* const err = sync + 3;
    gmem<double>* __restrict__ W = G(PD.W);
    const gmem<double>* __restrict__ b = G(PD.nb);
    gmem<double>* __restrict__ ny = G(PD.ny);
    gmem<double>* __restrict__ ninv = G(PD.ninv);
    gmem<double>* __restrict__ nxt = G(PD.nxt);
    extern __shared__ __align__(16) double lds[];
    double* Col = lds;  // nrows x kNewtonPS
    __shared__ double Dg[kNB][kNB + 1];   // the diagonal block -> L11 (lower)
    __shared__ double Xt[kNB][kNB + 1];   // L11^-T (upper)
    __shared__ double Dq[48][17];         // the diagonal factor's quarter products (rows 16..63)
    __shared__ double invd[kNB], ry[kNB];
    __shared__ int s_epoch, s_fail, s_last;
    __shared__ double s_sink;
    __shared__ int s_l11, s_rows, s_pub, s_upd;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int wave = tid >> 6, lane = tid & 63, nwaves = nthr >> 6;
    if (tid == 0) {
        s_epoch = __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
        s_fail = 0;
        s_l11 = 0;
        s_rows = 0;
        s_pub = 0;
        s_upd = 0;
    }
    for (int e = tid; e < kNB * kNB; e += nthr)
        Xt[e >> 5][e & (kNB - 1)] = ((e >> 5) == (e & (kNB - 1))) ? 1.0 : 0.0;  // the factor's identity rows
    const double shift = *PD.shift;
    // the block column's rows (W as the previous kernel left it), shift on
    // the diagonal. Only the lower d x d part of W belongs to this solve: the
    // padding rows and columns are formed here as the identity, and the
    // entries above the diagonal of the diagonal block start at zero (a
    // failed factorisation or the fallback solve may have left anything
    // there)
    // one row per thread (nrows <= kSegRows = blockDim.x); the row's 32
    // loads are all in flight before the first LDS store
    if (tid < nrows) {
        const int row = r0 + tid;
        double v[kNB];
#pragma unroll
        for (int c = 0; c < kNB; ++c) {
            const int col = kNB * J + c;
            v[c] = (row < d && col < d && row >= col) ? W[(size_t)col * dp + row] : 0.0;
        }
#pragma unroll
        for (int c = 0; c < kNB; ++c) {
            const int col = kNB * J + c;
            double t;
            if (row >= d || col >= d)
                t = (row == col) ? 1.0 : 0.0;
            else
                t = (row == col) ? v[c] + shift : v[c];
            Col[tid * kNewtonPS + c] = t;
        }
    }
    if (s == 0 && tid < kNB)
        ry[tid] = (kNB * J + tid < d) ? b[kNB * J + tid] : 0.0;
    __syncthreads();
    const int epoch = s_epoch;
    if (s == 0) PROBE_STAMP_AT(4000 + 16 * J + 0, 0);

    // 1. panels K < J; for (J, 0) the last one (K = J - 1) is on the critical
    //    path and is split below
    const int kbulk = (s == 0) ? J - 1 : J;
    for (int K = 0; K < kbulk; ++K) {
        if (tid < 64)
            mc_wait_ge(cnt + K, epoch * mc_nseg(dp, K), err);
        __syncthreads();
        mc_update_tiles(W, dp, K, r0, kNB * J, Col, wave, nwaves, ntile, s == 0, lane);
        if (s == 0 && tid < kNB) {
            // y_J -= P_K[block J] y_K
            const gmem<double>* pk = W + (size_t)(kNB * K) * dp + kNB * J + tid;
            double pv[kNB], yk[kNB];
#pragma unroll
            for (int k = 0; k < kNB; ++k) {
                pv[k] = ld_dev(pk + (size_t)k * dp);
                yk[k] = ld_dev(ny + kNB * K + k);
            }
            double acc = ry[tid];
#pragma unroll
            for (int k = 0; k < kNB; ++k)
                acc = fma(-pv[k], yk[k], acc);
            ry[tid] = acc;
        }
        __syncthreads();
    }

    // W at this segment's first row of block column J (solved rows go there)
    gmem<double>* __restrict__ Wcol = W + (size_t)(kNB * J) * dp + r0;
    if (s == 0) {
        const int K = J - 1;
        // published with the diagonal factor (dflag): L11, y_J, L11^-T and
        // the local row tiles 2 .. tpub - 1 (row blocks J+1 and J+2), so
        // that (J+1, 0) needs nothing else for its diagonal block and its
        // next row block
        const int tpub = min(ntile, 6);
        // L11^-T goes with the flag when lower segments wait for it
        const bool x_first = mc_nseg(dp, J) > 1;
        const int npub = 1 + (tpub > 3 ? 1 : 0) + (tpub > 4 ? 1 : 0) + (tpub > 5 ? 1 : 0) + (x_first ? 1 : 0);
        PROBE_STAMP_AT(4000 + 16 * J + 1, 0);
        if (wave == 0) {
            if (K >= 0) {
                mc_wait_ge(dflag + K, epoch, err);
                PROBE_STAMP_AT(4000 + 16 * J + 2, 0);
                mc_update_diag(W, ny, dp, K, kNB * J, Col, ry, lane);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            PROBE_STAMP_AT(4000 + 16 * J + 3, 0);
            if (lane < kNB) {
#pragma unroll
                for (int c = 0; c < kNB; ++c)
                    Dg[lane][c] = Col[lane * kNewtonPS + c];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // LDS scratch of the factor: the column broadcast in Dq (free
            // until the halves' MFMA step, and again after it), the per-column
            // records in Dg (read into registers before Dg is written)
            mc_diag_factor(Dg, Xt, Dq, invd, ry, &Dq[0][0], &Dg[0][0], lane, &s_fail);
            PROBE_STAMP_AT(4000 + 16 * J + 4, 0);
            if (lane == 0)
                mc_lds_signal(&s_l11, 1);
            // y_J out (L11 itself is read by nobody: the rows below use
            // L11^-T, which wave 4 stores, and the next round rebuilds W)
            if (lane < kNB)
                st_dev(ny + kNB * J + lane, ry[lane]);
            PROBE_STAMP_AT(4000 + 16 * J + 10, 0);
            // the first half of block J+1's rows
            if (tpub > 2) {
                mc_lds_wait(&s_rows, 1);
                PROBE_STAMP_AT(4000 + 16 * J + 11, 0);
                mc_solve_tiles(Col, Xt, 2, 1, 3, lane, Wcol, dp);
                PROBE_STAMP_AT(4000 + 16 * J + 12, 0);
            }
            if (s_fail && lane == 0)
                __hip_atomic_store(sync + 2, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            PROBE_STAMP_AT(4000 + 16 * J + 5, 0);
            // every publishing wave's stores have completed: the flag
            mc_drain();
            PROBE_STAMP_AT(4000 + 16 * J + 13, 0);
            if (lane == 0) {
                atomicAdd(&s_pub, 1);
                mc_lds_wait(&s_pub, npub);
                __hip_atomic_store(dflag + J, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            PROBE_STAMP_AT(4000 + 16 * J + 6, 0);
        } else {
            // waves 1..7. Wave 1 brings block J+1's rows (tiles 2, 3) up to
            // date with panel K (published with K's diagonal factor); wave 2
            // solves tile 3, waves 3 and 5 update and solve tiles 4 and 5
            // (block J+2: panel K's rows there come with K's whole column);
            // waves 1, 2, 3, 5, 6, 7 update tiles 6.. (wave 4 shares wave 0's
            // SIMD and issues no FP64 MFMA while the factor runs)
            if (wave == 1) {
                if (K >= 0 && tpub > 2) {
                    mc_wait_ge(dflag + K, epoch, err);
                    PROBE_STAMP_AT(4000 + 16 * J + 8, 64);
                    mc_update_tiles(W, dp, K, r0, kNB * J, Col, 2, 1, 4, false, lane);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                if (lane == 0)
                    mc_lds_signal(&s_rows, 1);
            }
            const bool mine45 = (wave == 3 && tpub > 4) || (wave == 5 && tpub > 5);
            const int t45 = (wave == 3) ? 4 : 5;
            const bool bulk = ntile > 6;
            if ((mine45 || (bulk && wave != 4)) && K >= 0)
                mc_wait_ge(cnt + K, epoch * mc_nseg(dp, K), err);
            if (mine45) {
                if (K >= 0)
                    mc_update_tiles(W, dp, K, r0, kNB * J, Col, t45, 1, t45 + 1, false, lane);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                mc_lds_wait(&s_l11, 1);
                mc_solve_tiles(Col, Xt, t45, 1, t45 + 1, lane, Wcol, dp);
                mc_drain();
                if (lane == 0)
                    atomicAdd(&s_pub, 1);
            }
            if (wave == 2 && tpub > 3) {
                mc_lds_wait(&s_l11, 1);
                mc_lds_wait(&s_rows, 1);
                mc_solve_tiles(Col, Xt, 3, 1, 4, lane, Wcol, dp);
                mc_drain();
                if (lane == 0)
                    atomicAdd(&s_pub, 1);
            }
            if (wave == 4) {
                // L11^-T by columns (nxt[J][c][q]): lane q + 32 h stores
                // columns 16 h .. 16 h + 15
                mc_lds_wait(&s_l11, 1);
                const int q = lane & (kNB - 1), h = lane >> 5;
                double v[kNB / 2];
#pragma unroll
                for (int c = 0; c < kNB / 2; ++c)
                    v[c] = Xt[q][16 * h + c];
#pragma unroll
                for (int c = 0; c < kNB / 2; ++c)
                    st_dev(nxt + (size_t)kNB * kNB * J + kNB * (16 * h + c) + q, v[c]);
                if (x_first) {
                    mc_drain();
                    if (lane == 0)
                        atomicAdd(&s_pub, 1);
                }
            }
            if (wave != 4) {
                const int bw = (wave <= 3) ? wave - 1 : wave - 2;  // 0..5 for waves 1, 2, 3, 5, 6, 7
                if (K >= 0 && bulk)
                    mc_update_tiles(W, dp, K, r0, kNB * J, Col, 6 + bw, 6, ntile, false, lane);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                if (lane == 0)
                    atomicAdd(&s_upd, 1);
            }
            PROBE_STAMP_AT(4000 + 16 * J + 9, 64);
        }
        // tiles 6.. by every wave once L11^-T and their updates are there
        if (ntile > 6) {
            mc_lds_wait(&s_l11, 1);
            mc_lds_wait(&s_upd, 6);
            mc_solve_tiles(Col, Xt, 6 + wave, nwaves, ntile, lane, Wcol, dp);
        }
    } else {
        // a lower segment: the triangular solve once L11^-T of block column J
        // is published (stored by columns: nxt[J][c][q])
        if (tid < 64)
            mc_wait_ge(dflag + J, epoch, err);
        __syncthreads();
        for (int e = tid; e < kNB * kNB; e += nthr)
            Xt[e & (kNB - 1)][e >> 5] = ld_dev(nxt + (size_t)kNB * kNB * J + e);
        __syncthreads();
        mc_solve_tiles(Col, Xt, wave, nwaves, ntile, lane, Wcol, dp);
    }
    // this segment of block column J is done (every wave's stores complete);
    // the ticket finds the last one
    mc_drain();
    __syncthreads();
    if (s == 0) PROBE_STAMP_AT(4000 + 16 * J + 7, 0);
    const int total = mc_blocks(dp);
    if (tid == 0) {
        __hip_atomic_fetch_add(cnt + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int t = __hip_atomic_fetch_add(sync + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == epoch * total - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!s_last)
        return;
    // every other workgroup's results are in memory: drop this CU's and this
    // XCD's stale cache lines, then plain (cached) loads
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

    // ---- the last workgroup: record, back solve (or fallback), iterate
    gmem<double>* __restrict__ x = G(PD.x);
    gmem<double>* __restrict__ dx_out = G(PD.dx);
    gmem<double>* __restrict__ trace = G(PD.trace);
    gmem<int>* __restrict__ round_ctr = G(PD.round_ctr);
    const int trace_cap = PD.trace_cap, update_x = PD.update_x;
    const int rnd = *round_ctr;
    if (PD.mst) {
        // an engine round: the record's ||grad f(x_r)|| (TrainLoop.h:1679-1691);
        // FedNL-PP reports the full gradient
        __shared__ double gsc[16];
        const double* __restrict__ gr = PD.pp ? PD.pp_tg : PD.nb;
        double gg = 0.0;
        for (int i = tid; i < d; i += nthr) gg = fma(gr[i], gr[i], gg);
        gg = block_sum(gg, gsc);
        if (tid == 0 && rnd < trace_cap)
            trace[(size_t)rnd * kTraceStride + 0] = sqrt(gg);
    }
    const bool failed = __hip_atomic_load(sync + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
    const bool expired = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    if (failed || expired) {
        if (failed && PD.H0) {
            __syncthreads();
            newton_fallback_solve(PD, lds);
        } else {
            for (int i = tid; i < d; i += nthr)
                dx_out[i] = 0.0;
        }
        if (tid == 0) {
            if (rnd < trace_cap) {
                trace[(size_t)rnd * kTraceStride + 3] = expired ? -1.0 : 0.0;
                trace[(size_t)rnd * kTraceStride + kTraceT1] = (double)wall_clock64();
            }
            *round_ctr = rnd + 1;
            __hip_atomic_store(sync, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    PROBE_STAMP_AT(4000 + 16 * 64 + 0, 0);
    // L^T dx = y by blocks from the bottom: dx_J = L11^-T r_J (wave 0, lane
    // q: row q of L11^-T), then every row rr above the block folds
    // r_rr -= sum_c L[32J + c][rr] dx_J[c], with L[32J + c][rr] =
    // W[rr dp + 32J + c]. Eight lanes share a row: lane u takes c = 4u ..
    // 4u + 3 (two 16-byte loads, so one load instruction covers eight rows of
    // 128 contiguous bytes) and the four products of the eight lanes are added
    // by a 3-step butterfly. Thread tid takes rows tid / 8 + 64 p, passes p in
    // chunks of kBsChunk; the next block's first chunk is in flight while the
    // current block runs. For dp <= 512 every block's L11^-T is staged in LDS
    // (row stride 33), and the L rows are first touched once per 128-byte
    // line, all loads in flight, so that the block loop reads them from this
    // XCD's L2.
    const bool staged = dp <= kSegRows;
    constexpr int kBsChunk = 8;
    const int rs = tid >> 3, u = tid & 7;
    const int rpp = nthr >> 3;  // rows per pass
    const gmem<dbl2>* __restrict__ W2 = (const gmem<dbl2>*)G(PD.W);
    double* Xs = lds;  // nb x 32 x 33 when staged
    double* yv = lds + (staged ? (size_t)nb * kNB * (kNB + 1) : 0);
    auto load_chunk = [&](int jb, int p0, dbl2 (&v)[kBsChunk][2]) {
#pragma unroll
        for (int p = 0; p < kBsChunk; ++p) {
            const int rr = rs + rpp * (p0 + p);
            if (rpp * (p0 + p) >= jb)
                break;  // no row of this pass (uniform)
            if (rr < jb) {
                const gmem<dbl2>* lp = W2 + ((size_t)rr * dp + jb + 4 * u) / 2;
                v[p][0] = lp[0];
                v[p][1] = lp[1];
            } else {
                v[p][0] = dbl2{0.0, 0.0};
                v[p][1] = dbl2{0.0, 0.0};
            }
        }
    };
    if (staged) {
        double warm = 0.0;
        if (tid < d) {
            const gmem<double>* __restrict__ wc = G(PD.W) + (size_t)tid * dp;
            const int rw = kNB * (tid / kNB + 1);
            double t[kSegRows / 16];
#pragma unroll
            for (int i = 0; i < kSegRows / 16; ++i) {
                const int r = rw + 16 * i;
                t[i] = (r < dp) ? wc[r] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < kSegRows / 16; ++i)
                warm += t[i];
        }
        if (d < 0)  // never: keeps the loads
            s_sink = warm;
        // L11^-T pairs (q, q + 1) of column c of block j: nxt[j][c][q]
        const gmem<dbl2>* __restrict__ x2 = (const gmem<dbl2>*)nxt;
        dbl2 xv[kSegRows / kNB];
#pragma unroll
        for (int k = 0; k < kSegRows / kNB; ++k) {
            const int e2 = tid + nthr * k;
            xv[k] = (e2 < nb * kNB * kNB / 2) ? x2[e2] : dbl2{0.0, 0.0};
        }
#pragma unroll
        for (int k = 0; k < kSegRows / kNB; ++k) {
            const int e2 = tid + nthr * k;
            if (e2 < nb * kNB * kNB / 2) {
                const int blkx = e2 >> 9, c = (e2 >> 4) & (kNB - 1), q = 2 * (e2 & 15);
                double* xb = Xs + (size_t)blkx * kNB * (kNB + 1);
                xb[q * (kNB + 1) + c] = xv[k].x;
                xb[(q + 1) * (kNB + 1) + c] = xv[k].y;
            }
        }
    }
    for (int i = tid; i < dp; i += nthr) yv[i] = ny[i];
    dbl2 vn[kBsChunk][2];
    load_chunk(kNB * (nb - 1), 0, vn);
    __syncthreads();
    PROBE_STAMP_AT(4000 + 16 * 64 + 2, 0);
    for (int Jb = nb - 1; Jb >= 0; --Jb) {
        const int jb = kNB * Jb;
        PROBE_STAMP_AT(4000 + 16 * 65 + Jb, 0);
        dbl2 v[kBsChunk][2];
#pragma unroll
        for (int p = 0; p < kBsChunk; ++p) {
            v[p][0] = vn[p][0];
            v[p][1] = vn[p][1];
        }
        if (Jb > 0)
            load_chunk(jb - kNB, 0, vn);
        if (wave == 0 && lane < kNB) {
            // four interleaved partial sums: a quarter of the dependent chain
            double a4[4] = {0.0, 0.0, 0.0, 0.0};
            if (staged) {
                const double* xr = Xs + (size_t)Jb * kNB * (kNB + 1) + lane * (kNB + 1);
#pragma unroll
                for (int c = 0; c < kNB; ++c)
                    a4[c & 3] = fma(xr[c], yv[jb + c], a4[c & 3]);
            } else {
                const gmem<double>* xr = nxt + (size_t)kNB * kNB * Jb + lane;  // X[lane][c] at xr[32 c]
#pragma unroll
                for (int c = 0; c < kNB; ++c)
                    a4[c & 3] = fma(xr[kNB * c], yv[jb + c], a4[c & 3]);
            }
            const double acc = (a4[0] + a4[1]) + (a4[2] + a4[3]);
            __builtin_amdgcn_wave_barrier();  // every lane has read r_J before any writes
            yv[jb + lane] = acc;
        }
        __syncthreads();
        const double dx0 = yv[jb + 4 * u], dx1 = yv[jb + 4 * u + 1];
        const double dx2 = yv[jb + 4 * u + 2], dx3 = yv[jb + 4 * u + 3];
        const int np = (jb + rpp - 1) / rpp;
        for (int p0 = 0; p0 < np; p0 += kBsChunk) {
            if (p0 > 0)
                load_chunk(jb, p0, v);
#pragma unroll
            for (int p = 0; p < kBsChunk; ++p) {
                if (p0 + p >= np)
                    break;
                const int rr = rs + rpp * (p0 + p);
                double t = v[p][0].x * dx0;
                t = fma(v[p][0].y, dx1, t);
                t = fma(v[p][1].x, dx2, t);
                t = fma(v[p][1].y, dx3, t);
                // the row's eight lanes: pairs, quads (DPP quad_perm), then
                // lane 8k takes lane 8k + 4 (row_shl:4)
                t += dpp_quad<0xB1>(t);
                t += dpp_quad<0x4E>(t);
                t += dpp_quad<0x104>(t);
                if (u == 0 && rr < jb)
                    yv[rr] = yv[rr] - t;
            }
        }
        __syncthreads();
    }
    PROBE_STAMP_AT(4000 + 16 * 64 + 1, 0);
    for (int i = tid; i < d; i += nthr) {
        const double v = yv[i];
        dx_out[i] = v;
        if (update_x == 1)
            x[i] = x[i] - v;
        else if (update_x == 2)
            x[i] = v;  // FedNL-PP: the solution is the next iterate
    }
    if (tid == 0) {
        if (rnd < trace_cap) {
            trace[(size_t)rnd * kTraceStride + 3] = 1.0;
            trace[(size_t)rnd * kTraceStride + kTraceT1] = (double)wall_clock64();
        }
        *round_ctr = rnd + 1;
        __hip_atomic_store(sync, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Lane layout checks of v_mfma_f32_16x16x4f32 and v_mfma_f32_16x16x32_bf16:
// in[0..63] / in[64..127] are the per-lane f32 operands, in[128..639] /
// in[640..1151] eight bf16 operands per lane (given as floats).
__global__ __launch_bounds__(64) void k_mfma_probe_mixed(const float* __restrict__ in, float* __restrict__ out)
{
    const int lane = threadIdx.x;
    const flt4 z = {0.f, 0.f, 0.f, 0.f};
    const flt4 r1 = __builtin_amdgcn_mfma_f32_16x16x4f32(in[lane], in[64 + lane], z, 0, 0, 0);
    bf16x8 a, b;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        a[u] = (__bf16)in[128 + lane * 8 + u];
        b[u] = (__bf16)in[640 + lane * 8 + u];
    }
    const flt4 r2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, z, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        out[lane * 4 + q] = r1[q];
        out[256 + lane * 4 + q] = r2[q];
    }
}

// ---------------------------------------------------------------------------
// host engine
// ---------------------------------------------------------------------------
enum KernelId {
    KID_MARGIN, KID_HESS, KID_SELECT, KID_EMIT, KID_MASTER, KID_NEWTON, KID_LS, KID_GRAD, KID_GD, KID_COUNT
};
static const char* kKernelNames[KID_COUNT] = {"k_margin", "k_hessian_f64", "k_select", "k_emit", "k_master",
                                              "k_newton_mc", "k_line_search", "k_grad_f64", "k_gd"};

struct Engine {
    fednl_opts o;
    int n = 0, m = 0, d = 0, k = 0, lda = 0, m_pad = 0, d_pad = 0, nt = 0, tiles_up = 0, dp = 0, workers = 1;
    int nparts = 0;  // Lk partial slots per client (tiles_up, or the sparse form's segments)
    size_t S = 0;
    double alpha = 1.0, prescale = 0.0, inv_n = 0.0, mult = 1.0, lek_delta = 0.0;
    hipStream_t st = nullptr;

    double *A = nullptr, *x = nullptr, *Dw = nullptr, *gw = nullptr, *nlog = nullptr;
    double *learned = nullptr, *delta = nullptr, *grad_c = nullptr, *part = nullptr, *lk = nullptr;
    uint32_t *sel = nullptr, *pos = nullptr, *idx = nullptr, *rng = nullptr, *lek_state = nullptr, *lek_popped = nullptr;
    uint32_t *seq_start = nullptr, *sel_bm = nullptr;
    int bm_words = 0;
    int *sel_cnt = nullptr, *cnt = nullptr, *lek_hit = nullptr, *round_ctr = nullptr;
    double *lek_p = nullptr, *vals = nullptr, *H = nullptr, *g_avg = nullptr, *shift = nullptr, *W = nullptr;
    double *dx = nullptr, *trace = nullptr;
    int* nsync = nullptr;   // multi-workgroup Newton solve: sync words
    double *ny = nullptr, *ninv = nullptr, *nxt = nullptr;
    double* mst = nullptr;  // master state: MstSlot header + x_k
    double* H0 = nullptr;   // H at creation (S, packed upper)
    // FedNL-PP state (clients_per_round > 0)
    int* pp_sel = nullptr;
    uint32_t* pp_rng = nullptr;
    double *pp_dir = nullptr, *pp_lk = nullptr, *pp_gmsg = nullptr, *pp_lkmsg = nullptr, *pp_h = nullptr;
    double* pp_tg = nullptr;
    int trace_cap = 0;
    int rounds_enqueued = 0;

    double wclk_hz = 1e8;  // device wall clock rate (the trace's round timestamps)
    bool timing = false;
    unsigned timing_mask = ~0u;
    // FEDNL_ROUND_GRAPH=1: one round captured into a HIP graph and replayed
    // (unless kernel timing is on). Off by default: on MI355X the replay
    // measured 3% below plain launches, which already run back to back
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int margin_rows = 4;           // k_margin rows per wave, picked at creation
    int hess_chunk = 16;           // k_hessian_f64 samples per LDS chunk, picked at creation
    struct Pending { int kid; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;
    double kms[KID_COUNT] = {0};
    int klaunch[KID_COUNT] = {0};
    size_t lds_select = 0, lds_select_b = 0, lds_emit = 0, lds_newton = 0;  // lds_select_b: as a batch member
    bool wide = false;             // selection heap / bitmap in global memory
    unsigned char* sel_scr = nullptr;  // wide selection: per-client heap or sort keys
    size_t sel_scr_stride = 0;
    int hmode = 0;                 // 0 FP64, 1 FP32, 2 BF16 split (A^T D A only)
    ProblemDev hdesc{};            // descriptor as last uploaded
    ProblemDev* pdev = nullptr;    // device copy read by every round kernel
    float* A32 = nullptr;          // FP32 copy of A (mode 1)
    __bf16* AThi = nullptr;        // transposed BF16 hi / lo of A (mode 2)
    __bf16* ATlo = nullptr;
    double* gpart = nullptr;       // modes 1 / 2 without FedNL-PP: margin-pass gradient partials
    int hform = FEDNL_HESS_DENSE;  // the FP64 Hessian form picked at creation
    uint4* sp_seg = nullptr;       // sparse form: segment records, lane tables, streams (k_hessian_sparse)
    uint32_t* sp_lane = nullptr;
    uint16_t* sp_stream = nullptr;
    double* sp_b = nullptr;
    uint32_t* sp_rowptr = nullptr;  // sparse form: the margin pass's row lists (k_margin_sparse)
    uint16_t* sp_cols = nullptr;
    size_t lds_sparse = 0;
    size_t sp_bytes = 0;           // device bytes of the sparse form's lists
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
        if (e->timing && ((e->timing_mask >> kid) & 1u)) {
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
    HIP_OK(hipMalloc(&nt, sizeof(double) * kTraceStride * (size_t)cap));
    HIP_OK(hipMemsetAsync(nt, 0, sizeof(double) * kTraceStride * (size_t)cap, e->st));
    if (e->trace) {
        HIP_OK(hipMemcpyAsync(nt, e->trace, sizeof(double) * kTraceStride * (size_t)e->trace_cap, hipMemcpyDeviceToDevice, e->st));
        HIP_OK(hipStreamSynchronize(e->st));
        HIP_OK(hipFree(e->trace));
    }
    e->trace = nt;
    e->trace_cap = cap;
    return FEDNL_OK;
}

// the engine's descriptor as the kernels see it
// The heap selection's LDS: the heap values and ids, then the scan's ring of
// staged groups (as many slots as keep the selection near 64 KB, so that two
// selections share a CU; at least 2, at most kMaxScanRing).
static size_t select_heap_bytes(int k, size_t S)
{
    const int P = pow2_at_least(k);
    const int ids = select_bitmap_fits(S, k) ? k : P;
    return (size_t)8 * (k + 4) + (size_t)4 * round_up(ids + 4, 2);
}
// LDS slots of the TopK scan's ring: up to 96 KB with the heap when every
// selection of the launch has a CU of its own (big), else up to 64 KB so
// that two selections share a CU
static int scan_ring_slots(int k, size_t S, bool big)
{
    const size_t slot = (size_t)8 * kScanGroup * 64;
    const size_t heap = select_heap_bytes(k, S);
    const size_t target = (big ? 96 : 64) * 1024;
    const int r = heap + 2 * slot >= target ? 2 : (int)((target - heap) / slot);
    return std::max(2, std::min(kMaxScanRing, r));
}
static size_t select_lds_bytes(int k, size_t S, bool big)
{
    const int P = pow2_at_least(k);
    const size_t heap = select_heap_bytes(k, S) + (size_t)8 * kScanGroup * 64 * scan_ring_slots(k, S, big);
    const size_t rk = (size_t)8 * P + (size_t)8 * k;
    return std::max(heap, rk);
}

// the compressor keeps a heap (TopK, TopLEK: unless k covers the whole
// list) or sorts k keys (RandK: always, k = S included)
static bool select_keeps_k(int comp, int k, size_t S)
{
    return ((comp == FEDNL_TOPK || comp == FEDNL_TOPLEK) && (size_t)k < S) || comp == FEDNL_RANDK;
}
// Wide selection: k above the LDS-resident heap, or a selection bitmap of
// S positions above kEmitLdsMax. The heap (or RandK's keys and chains) then
// lives in per-client global scratch and the bitmap is built in global memory.
static bool select_wide(int comp, int k, size_t S)
{
    return (select_keeps_k(comp, k, S) && k > kMaxHeapK) || ((S + 31) / 32) * 4 > (size_t)kEmitLdsMax;
}
// bytes of wide-selection scratch per client (16-byte multiple)
static size_t select_scratch_stride(int comp, int k, size_t S)
{
    if (!select_keeps_k(comp, k, S))
        return 0;
    const size_t P = (size_t)pow2_at_least(k);
    const size_t heap = (size_t)12 * (k + 4);
    const size_t rk = (size_t)8 * P + (size_t)8 * k;
    return (std::max(heap, rk) + 15) / 16 * 16;
}
// dynamic LDS of the selection proper (the fused emit adds its bitmap)
static size_t select_lds(int comp, int k, size_t S, bool wide, bool big)
{
    if (!select_keeps_k(comp, k, S))
        return 0;
    if (!wide)
        return select_lds_bytes(k, S, big);
    if (comp == FEDNL_RANDK)
        return 0;
    // the scan's ring, later the tail's sort of the ids
    const size_t ring = (size_t)8 * kScanGroup * 64 * kMaxScanRing;
    return std::max(ring, (size_t)4 * pow2_at_least(k));
}
static size_t emit_lds(size_t S, bool wide) { return wide ? 0 : sizeof(uint32_t) * ((S + 31) / 32); }

// the selection of a solo engine with at most kBigRingClients clients runs
// one workgroup per CU with the big scan ring; batch members use the small one
constexpr int kBigRingClients = 256;
static bool big_ring(const Engine* e, bool batched) { return !batched && e->n <= kBigRingClients; }

static ProblemDev make_desc(const Engine* e, bool batched = false)
{
    ProblemDev P{};
    P.n = e->n; P.m = e->m; P.d = e->d; P.k = e->k; P.lda = e->lda; P.m_pad = e->m_pad; P.d_pad = e->d_pad;
    P.tiles_up = e->tiles_up; P.nparts = e->nparts; P.dp = e->dp; P.workers = e->workers; P.comp = e->o.compressor; P.hmode = e->hmode;
    P.option_b = e->o.option_b; P.tracking = e->o.tracking; P.bm_words = e->bm_words;
    P.trace_cap = e->trace_cap; P.S = e->S;
    P.pp = e->o.clients_per_round > 0 ? 1 : 0; P.pp_tau = e->o.clients_per_round;
    P.gd_step = e->o.global_lr;
    // x -= dx (FedNL), x = solution (FedNL-PP), x left to the line search (FedNL-LS)
    P.update_x = P.pp ? 2 : (e->o.line_search ? 0 : 1);
    P.ls = e->o.line_search ? 1 : 0; P.ls_max = e->o.max_line_search_steps;
    P.ls_c = e->o.c_line_search; P.ls_gamma = e->o.gamma_line_search; P.mst = e->mst; P.H0 = e->H0;
    P.pp_sel = e->pp_sel; P.pp_rng = e->pp_rng; P.pp_dir = e->pp_dir; P.pp_lk = e->pp_lk; P.pp_gmsg = e->pp_gmsg;
    P.pp_lkmsg = e->pp_lkmsg; P.pp_h = e->pp_h; P.pp_tg = e->pp_tg;
    P.lambda = e->o.lambda; P.alpha = e->alpha; P.prescale = e->prescale; P.inv_n = e->inv_n; P.mult = e->mult;
    P.lek_delta = e->lek_delta;
    P.A = e->A; P.A32 = e->A32; P.AThi = e->AThi; P.ATlo = e->ATlo;
    P.x = e->x; P.Dw = e->Dw; P.gw = e->gw; P.nlog = e->nlog; P.learned = e->learned; P.delta = e->delta;
    P.gpart = e->gpart;
    P.sp_seg = e->sp_seg; P.sp_lane = e->sp_lane; P.sp_stream = e->sp_stream; P.sp_b = e->sp_b;
    P.sp_rowptr = e->sp_rowptr; P.sp_cols = e->sp_cols;
    P.grad_c = e->grad_c; P.part = e->part; P.lk = e->lk; P.vals = e->vals; P.H = e->H; P.g_avg = e->g_avg;
    P.shift = e->shift; P.W = e->W; P.dx = e->dx; P.trace = e->trace; P.lek_p = e->lek_p;
    P.hup = nullptr; P.nb = e->g_avg;
    P.sel = e->sel; P.pos = e->pos; P.idx = e->idx; P.rng = e->rng; P.lek_state = e->lek_state;
    P.lek_popped = e->lek_popped; P.seq_start = e->seq_start; P.sel_bm = e->sel_bm;
    P.sel_cnt = e->sel_cnt; P.cnt = e->cnt; P.lek_hit = e->lek_hit; P.round_ctr = e->round_ctr;
    P.nsync = e->nsync; P.ny = e->ny; P.ninv = e->ninv; P.nxt = e->nxt;
    P.sel_ring = e->wide ? kMaxScanRing : scan_ring_slots(std::min<int>(e->k, kMaxHeapK), e->S, big_ring(e, batched));
    P.wide = e->wide ? 1 : 0;
    P.sel_scr = e->sel_scr;
    P.sel_scr_stride = e->sel_scr_stride;
    return P;
}

// uploads the descriptor when a field changed (e.g. the trace buffer grew)
static int sync_desc(Engine* e)
{
    const ProblemDev P = make_desc(e);
    if (memcmp(&P, &e->hdesc, sizeof(P)) == 0)
        return FEDNL_OK;
    e->hdesc = P;
    HIP_OK(hipMemcpyAsync(e->pdev, &e->hdesc, sizeof(ProblemDev), hipMemcpyHostToDevice, e->st));
    // the source is a member that the next call may change: finish the copy now
    HIP_OK(hipStreamSynchronize(e->st));
    return FEDNL_OK;
}

// number of blocks each round kernel uses for one problem
static int blocks_margin(const ProblemDev& P) { return (P.m_pad / 16) * P.n; }
// dynamic LDS of k_margin<false, 4, true>: x and the four waves' column sums
static size_t margin_grad_lds(int d_pad) { return 5 * sizeof(double) * (size_t)d_pad; }

// the engine's margin pass with its tuned rows per wave
static int blocks_margin_sparse(const ProblemDev& P) { return ((P.m_pad + kSpMarginRows - 1) / kSpMarginRows) * P.n; }

static void launch_margin(Engine* e, bool trial, const ProblemDev* pd_over = nullptr)
{
    const ProblemDev& P = e->hdesc;
    if (e->hform == FEDNL_HESS_SPARSE) {
        const ProblemDev* q = pd_over ? pd_over : e->pdev;
        const int2* nomap = nullptr;
        const size_t lds = sizeof(double) * e->d_pad;
        if (trial)
            hipLaunchKernelGGL(k_margin_sparse<true>, dim3(blocks_margin_sparse(P)), dim3(kSpMarginRows), lds, e->st,
                               q, nomap);
        else
            hipLaunchKernelGGL(k_margin_sparse<false>, dim3(blocks_margin_sparse(P)), dim3(kSpMarginRows), lds, e->st,
                               q, nomap);
        return;
    }
    const unsigned grid = (unsigned)((P.m_pad / (4 * e->margin_rows)) * P.n);
    const size_t lds = sizeof(double) * e->d_pad;
    const ProblemDev* pd = e->pdev;
    const int2* nomap = nullptr;
    if (e->gpart && !trial) {
        // with the client gradients' partials (16 rows per workgroup)
        hipLaunchKernelGGL((k_margin<false, 4, true>), dim3(grid), dim3(256), margin_grad_lds(e->d_pad), e->st, pd,
                           nomap);
        return;
    }
    if (e->margin_rows == 8) {
        if (trial)
            hipLaunchKernelGGL((k_margin<true, 8>), dim3(grid), dim3(256), lds, e->st, pd, nomap);
        else
            hipLaunchKernelGGL((k_margin<false, 8>), dim3(grid), dim3(256), lds, e->st, pd, nomap);
    } else {
        if (trial)
            hipLaunchKernelGGL((k_margin<true, 4>), dim3(grid), dim3(256), lds, e->st, pd, nomap);
        else
            hipLaunchKernelGGL((k_margin<false, 4>), dim3(grid), dim3(256), lds, e->st, pd, nomap);
    }
}

// Tile autotuning at engine creation: each candidate shape of the margin
// pass (4 or 8 rows per wave) runs on the engine's own data, timed with HIP
// events on its stream after a warm-up, and the fastest is kept. Both shapes
// give the same bits (same per-row sums, -log sigma partials per 16 rows).
// FEDNL_MARGIN_ROWS=4 or 8 sets the shape instead.
static void tune_margin(Engine* e)
{
    if (e->hform == FEDNL_HESS_SPARSE)
        return;  // the sparse form's margin pass has one shape
    if (e->gpart) {
        // the gradient partials are per 16 rows: four rows per wave
        e->margin_rows = 4;
        return;
    }
    const char* env = getenv("FEDNL_MARGIN_ROWS");
    if (env && (env[0] == '4' || env[0] == '8') && env[1] == 0) {
        e->margin_rows = env[0] - '0';
        return;
    }
    static const int cand[2] = {4, 8};
    float best = 0.f;
    int pick = 4;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        if (a) hipEventDestroy(a);
        e->margin_rows = 4;
        return;
    }
    for (int q = 0; q < 2; ++q) {
        e->margin_rows = cand[q];
        for (int w = 0; w < 3; ++w) launch_margin(e, false);
        hipEventRecord(a, e->st);
        for (int r = 0; r < 10; ++r) launch_margin(e, false);
        hipEventRecord(b, e->st);
        float ms = 0.f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
            ms = 1e30f;
        if (q == 0 || ms < best) {
            best = ms;
            pick = cand[q];
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    e->margin_rows = pick;
}
static int blocks_hessian(const ProblemDev& P) { return 8 * ((P.n + 7) / 8) * P.tiles_up; }
static int blocks_sparse(const ProblemDev& P) { return 8 * ((P.n + 7) / 8) * P.nparts; }
static int blocks_grad(const ProblemDev& P) { return (P.d_pad / kTile) * P.n; }
static int blocks_master(const ProblemDev& P) { return master_h_blocks(P.S) + master_grad_blocks(P.d) + 1; }

// client Hessians (and, for the reduced-precision modes, the FP64 client
// gradients) at the current margin weights; init: write learned (or
// out_learned) instead of Delta. pd / map / nblk: a descriptor array and
// block map (batches); null pd means the engine itself.
static void launch_hessian_on(hipStream_t st, int hmode, bool init, const ProblemDev* pd, const int2* map,
                              int nblk_h, int nblk_g, double* out_learned, bool fused_grad, int hchunk = 16)
{
    if (hmode == 0) {
        if (init)
            hipLaunchKernelGGL((k_hessian_f64<true, 16>), dim3(nblk_h), dim3(256), 0, st, pd, map, out_learned);
        else if (hchunk == 32)
            hipLaunchKernelGGL((k_hessian_f64<false, 32>), dim3(nblk_h), dim3(256), 0, st, pd, map, out_learned);
        else
            hipLaunchKernelGGL((k_hessian_f64<false, 16>), dim3(nblk_h), dim3(256), 0, st, pd, map, out_learned);
        return;
    }
    if (!init && !fused_grad)
        hipLaunchKernelGGL(k_grad_f64, dim3(nblk_g), dim3(256), 0, st, pd, map == nullptr ? nullptr : map + nblk_h);
    if (hmode == 1) {
        if (init)
            hipLaunchKernelGGL(k_hessian_f32<true>, dim3(nblk_h), dim3(256), 0, st, pd, map, out_learned);
        else
            hipLaunchKernelGGL(k_hessian_f32<false>, dim3(nblk_h), dim3(256), 0, st, pd, map, out_learned);
    } else {
        if (init)
            hipLaunchKernelGGL(k_hessian_bf16<true>, dim3(nblk_h), dim3(256), 0, st, pd, map, out_learned);
        else
            hipLaunchKernelGGL(k_hessian_bf16<false>, dim3(nblk_h), dim3(256), 0, st, pd, map, out_learned);
    }
}

static void launch_hessian(Engine* e, bool init, double* out_learned, const ProblemDev* pd = nullptr)
{
    const ProblemDev& P = e->hdesc;
    if (e->hform == FEDNL_HESS_SPARSE) {
        const ProblemDev* q = pd ? pd : e->pdev;
        const int2* nomap = nullptr;
        if (init)
            hipLaunchKernelGGL(k_hessian_sparse<true>, dim3(blocks_sparse(P)), dim3(kSpThreads), e->lds_sparse, e->st, q,
                               nomap, out_learned);
        else
            hipLaunchKernelGGL(k_hessian_sparse<false>, dim3(blocks_sparse(P)), dim3(kSpThreads), e->lds_sparse, e->st, q,
                               nomap, out_learned);
        return;
    }
    launch_hessian_on(e->st, e->hmode, init, pd ? pd : e->pdev, nullptr, blocks_hessian(P), blocks_grad(P),
                      out_learned, e->gpart != nullptr, e->hess_chunk);
}

// The FP64 Hessian's chunk depth (16 or 32 samples per LDS stage), timed at
// creation like the margin pass: each candidate on the engine's own data
// (after the learned Hessians are set; a round's Hessian pass at x0 only
// rewrites Delta, the Lk partials and the client gradients, which round 0
// recomputes), five launches each after a warm-up. Both give the same bits.
// FEDNL_HESS_CHUNK=16 or 32 sets the depth instead. (Measured on MI355X the
// 16-deep stage wins at every shape tried: the 32-deep one holds two
// workgroups per CU for its 80 KB of LDS.)
static void tune_hessian(Engine* e)
{
    if (e->hmode != 0 || e->hform != FEDNL_HESS_DENSE)
        return;
    const char* env = getenv("FEDNL_HESS_CHUNK");
    if (env && (!strcmp(env, "16") || !strcmp(env, "32"))) {
        e->hess_chunk = atoi(env);
        return;
    }
    static const int cand[2] = {16, 32};
    float best = 0.f;
    int pick = 16;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        if (a) hipEventDestroy(a);
        e->hess_chunk = 16;
        return;
    }
    for (int q = 0; q < 2; ++q) {
        e->hess_chunk = cand[q];
        for (int w = 0; w < 2; ++w) launch_hessian(e, false, nullptr);
        hipEventRecord(a, e->st);
        for (int r = 0; r < 5; ++r) launch_hessian(e, false, nullptr);
        hipEventRecord(b, e->st);
        float ms = 0.f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
            ms = 1e30f;
        if (q == 0 || ms < best) {
            best = ms;
            pick = cand[q];
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    e->hess_chunk = pick;
}

// one GD round: margins (gradient weights, f partials), client gradients,
// their average with f and ||x||, the step
static int launch_round_gd(Engine* e)
{
    const ProblemDev& P = e->hdesc;
    const ProblemDev* pd = e->pdev;
    const int2* nomap = nullptr;
    {
        KTime kt(e, KID_MARGIN);
        launch_margin(e, false);
    }
    {
        KTime kt(e, KID_GRAD);
        hipLaunchKernelGGL(k_grad_f64, dim3(blocks_grad(P)), dim3(256), 0, e->st, pd, nomap);
    }
    {
        KTime kt(e, KID_GD);
        hipLaunchKernelGGL(k_gd_master, dim3(master_grad_blocks(P.d) + 1), dim3(256), 0, e->st, pd, nomap);
        hipLaunchKernelGGL(k_gd_step, dim3(1), dim3(256), 0, e->st, pd, nomap);
    }
    return FEDNL_OK;
}

static int launch_round(Engine* e)
{
    if (e->o.algorithm == FEDNL_ALG_GD)
        return launch_round_gd(e);
    const ProblemDev& P = e->hdesc;
    const ProblemDev* pd = e->pdev;
    const int2* nomap = nullptr;
    const int n = e->n;
    const bool pp = e->pp_sel != nullptr;
    if (pp)
        hipLaunchKernelGGL(k_pp_sample, dim3(1), dim3(64), 0, e->st, pd, nomap);
    {
        KTime kt(e, KID_MARGIN);
        launch_margin(e, false);
    }
    {
        KTime kt(e, KID_HESS);
        // FedNL-PP reports the full gradient: every client's, from k_grad_f64
        // (the reduced-precision modes launch it themselves)
        if (pp && e->hmode == 0)
            hipLaunchKernelGGL(k_grad_f64, dim3(blocks_grad(P)), dim3(256), 0, e->st, pd, nomap);
        launch_hessian(e, false, nullptr);
    }
    const int comp = e->o.compressor;
    {
        KTime kt(e, KID_SELECT);
        // the messages follow the selection in the same workgroup, except
        // for TopLEK (its draw runs over the other clients' outcomes)
        const size_t lds_bm = e->lds_emit;
        const size_t lds_fused = std::max(e->lds_select, lds_bm);
        const dim3 g(n), t(256);
        if (!e->wide) {
            if (comp == FEDNL_TOPK)
                hipLaunchKernelGGL((k_select_heap<false, true, false>), g, t, lds_fused, e->st, pd, nomap);
            else if (comp == FEDNL_TOPLEK)
                hipLaunchKernelGGL((k_select_heap<true, false, false>), g, t, e->lds_select, e->st, pd, nomap);
            else if (comp == FEDNL_RANDK)
                hipLaunchKernelGGL((k_select_randk<true, false>), g, t, lds_fused, e->st, pd, nomap);
            else
                hipLaunchKernelGGL((k_select_window<true, false>), g, t, lds_bm, e->st, pd, nomap);
        } else {
            if (comp == FEDNL_TOPK)
                hipLaunchKernelGGL((k_select_heap<false, true, true>), g, t, lds_fused, e->st, pd, nomap);
            else if (comp == FEDNL_TOPLEK)
                hipLaunchKernelGGL((k_select_heap<true, false, true>), g, t, e->lds_select, e->st, pd, nomap);
            else if (comp == FEDNL_RANDK)
                hipLaunchKernelGGL((k_select_randk<true, true>), g, t, lds_fused, e->st, pd, nomap);
            else
                hipLaunchKernelGGL((k_select_window<true, true>), g, t, lds_bm, e->st, pd, nomap);
        }
    }
    if (comp == FEDNL_TOPLEK || pp) {
        KTime kt(e, KID_EMIT);
        if (comp == FEDNL_TOPLEK) {
            if (e->wide)
                hipLaunchKernelGGL(k_emit<true>, dim3(n), dim3(256), 0, e->st, pd, nomap);
            else
                hipLaunchKernelGGL(k_emit<false>, dim3(n), dim3(256), e->lds_emit, e->st, pd, nomap);
        }
        if (pp)
            hipLaunchKernelGGL(k_pp_client<false>, dim3(n), dim3(256), 0, e->st, pd, nomap);
    }
    {
        KTime kt(e, KID_MASTER);
        hipLaunchKernelGGL(k_master, dim3(blocks_master(P)), dim3(256), 0, e->st, pd, nomap);
    }
    {
        KTime kt(e, KID_NEWTON);
        hipLaunchKernelGGL(k_newton_mc, dim3(mc_blocks(P.dp)), dim3(kMcThreads), e->lds_newton, e->st, pd, nomap);
    }
    if (e->o.line_search) {
        // every trial is enqueued; the ones after acceptance return at once
        KTime kt(e, KID_LS);
        hipLaunchKernelGGL(k_ls_begin, dim3(1), dim3(256), 0, e->st, pd, nomap);
        for (int t = 0; t < e->o.max_line_search_steps; ++t) {
            launch_margin(e, true);
            hipLaunchKernelGGL(k_ls_check, dim3(1), dim3(256), 0, e->st, pd, nomap);
        }
    }
    return FEDNL_OK;
}


static void free_engine(Engine* e)
{
    if (!e)
        return;
    if (e->gexec)
        hipGraphExecDestroy(e->gexec);
    if (e->graph)
        hipGraphDestroy(e->graph);
    void* ptrs[] = {e->A, e->x, e->Dw, e->gw, e->nlog, e->learned, e->delta, e->grad_c, e->part, e->lk, e->sel_bm,
                    e->sel, e->pos, e->idx, e->rng, e->lek_state, e->lek_popped, e->seq_start, e->sel_cnt,
                    e->cnt, e->lek_hit, e->round_ctr, e->lek_p, e->vals, e->H, e->g_avg, e->shift, e->W,
                    e->dx, e->trace, e->A32, e->AThi, e->ATlo, e->pdev, e->mst, e->H0, e->pp_sel, e->pp_rng,
                    e->pp_dir, e->pp_lk, e->pp_gmsg, e->pp_lkmsg, e->pp_h, e->pp_tg, e->nsync, e->ny, e->ninv,
                    e->nxt, e->sel_scr, e->gpart, e->sp_seg, e->sp_lane, e->sp_stream, e->sp_b,
                    e->sp_rowptr, e->sp_cols};
    for (void* p : ptrs)
        if (p)
            hipFree(p);
    for (auto& p : e->pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (auto v : e->ev_pool) hipEventDestroy(v);
    if (e->st)
        hipStreamDestroy(e->st);
    delete e;
}

template <class T>
static hipError_t dalloc(T** p, size_t count)
{
    return hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
}

// Raises a kernel's dynamic-LDS limit to at least `bytes` (never lowers it:
// engines of other sizes may still launch the kernel with their own larger
// size). The limit is per device.
static hipError_t raise_lds_limit(const void* func, size_t bytes)
{
    if (bytes <= 64 * 1024)
        return hipSuccess;
    static std::mutex mu;
    static std::vector<std::pair<std::pair<int, const void*>, size_t>> set;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lk(mu);
    for (auto& s : set)
        if (s.first.first == dev && s.first.second == func) {
            if (s.second >= bytes)
                return hipSuccess;
            e = hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e == hipSuccess)
                s.second = bytes;
            return e;
        }
    e = hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess)
        set.push_back({{dev, func}, bytes});
    return e;
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

// ---------------------------------------------------------------------------
// Sparse Hessian form (k_hessian_sparse): eligibility and the token streams
// ---------------------------------------------------------------------------
static bool design_is_binary(const double* f, size_t count)
{
    for (size_t i = 0; i < count; ++i)
        if (f[i] != 0.0 && f[i] != 1.0)
            return false;
    return true;
}

// sum over samples of z (z + 1) / 2, z the sample's nonzero features: the
// terms the sparse form adds
static double sparse_pair_count(const double* f, size_t rows, int d)
{
    double tot = 0.0;
    for (size_t r = 0; r < rows; ++r) {
        const double* a = f + r * d;
        size_t z = 0;
        for (int i = 0; i < d; ++i) z += (a[i] != 0.0);
        tot += 0.5 * (double)z * (double)(z + 1);
    }
    return tot;
}

// an item's cost to its lane in k_hessian_sparse, in dependent LDS round
// trips: header (with four samples for a short off-diagonal entry) and
// weights; else header, batches of sixteen with the next batch's indices read
// ahead, and the rest in one masked batch
static uint32_t sparse_item_cost(uint32_t cnt, bool short_entry)
{
    if (short_entry)
        return 2;
    return 2 + (cnt >= 16 ? 1 + (cnt >> 4) : 0) + ((cnt & 15u) ? 2 : 0);
}

// an item of at least 32 samples that are consecutive goes as a run
static bool sparse_is_run(const uint16_t* h, uint32_t k)
{
    if (k < 32)
        return false;
    return (uint32_t)h[k - 1] - h[0] == k - 1;  // ascending and distinct
}

struct SparsePlan {
    std::vector<uint32_t> seg;     // per (client, dispatch rank): segment, length, stream offset (lo, hi)
    std::vector<uint32_t> lane;    // per (client, dispatch rank): kSpLanes + 1 lane offsets
    std::vector<uint16_t> stream;  // all token streams
    std::vector<double> b;         // per (client, padded sample): label sign, 0 on padding
    std::vector<uint32_t> rowptr;  // per client m + 1 offsets into cols
    std::vector<uint16_t> cols;    // every sample's nonzero columns, ascending
    size_t max_tokens = 0;         // longest segment stream
};

// One client's part of the plan (its segments in segment order).
struct SparseClient {
    std::vector<std::vector<uint16_t>> tok;  // per segment: the lanes' tokens
    std::vector<std::vector<uint32_t>> lo;   // per segment: kSpLanes + 1 lane offsets
    std::vector<uint64_t> mx;                // per segment: its costliest lane
    std::vector<uint32_t> rowptr;            // m + 1 offsets into cols
    std::vector<uint16_t> cols;              // the samples' nonzero columns
};

// For client c: the samples of each upper entry (ascending); per segment of
// kSpSeg entries the items (listed entries, and the gradient of each column
// whose diagonal entry the segment holds) spread over kSpLanes lanes,
// costliest first, each to the lane with the least cost so far (ties in an
// order that alternates the waves), and the lanes' tokens in lane order.
static void build_sparse_client(const double* features, int c, int m, int d, const std::vector<int>& diag_col,
                                SparseClient& out)
{
    const size_t S = (size_t)d * (d + 1) / 2;
    const int nseg = (int)((S + kSpSeg - 1) / kSpSeg);
    constexpr int kW = kSpLanes / 64;
    std::vector<uint32_t> cnt(S, 0u), start(S + 1), cur(S);
    std::vector<uint16_t> hits;
    std::vector<uint32_t>& rowptr = out.rowptr;
    std::vector<uint16_t>& cols = out.cols;
    rowptr.assign(m + 1, 0);
    cols.clear();
    struct Item {
        uint32_t cost, e;
        bool grad;
    };
    std::vector<Item> items;
    std::vector<std::vector<Item>> lane_items(kSpLanes);
    using Load = std::pair<uint64_t, int>;
    for (int s = 0; s < m; ++s) {
        rowptr[s] = (uint32_t)cols.size();
        const double* a = features + ((size_t)c * m + s) * d;
        for (int i = 0; i < d; ++i)
            if (a[i] != 0.0)
                cols.push_back((uint16_t)i);
    }
    rowptr[m] = (uint32_t)cols.size();
    for (int s = 0; s < m; ++s)
        for (uint32_t bq = rowptr[s]; bq < rowptr[s + 1]; ++bq) {
            const size_t col0 = (size_t)cols[bq] * (cols[bq] + 1) / 2;
            for (uint32_t aq = rowptr[s]; aq <= bq; ++aq) cnt[col0 + cols[aq]]++;
        }
    start[0] = 0;
    for (size_t e = 0; e < S; ++e) start[e + 1] = start[e] + cnt[e];
    hits.resize(start[S]);
    std::copy(start.begin(), start.begin() + S, cur.begin());
    for (int s = 0; s < m; ++s)
        for (uint32_t bq = rowptr[s]; bq < rowptr[s + 1]; ++bq) {
            const size_t col0 = (size_t)cols[bq] * (cols[bq] + 1) / 2;
            for (uint32_t aq = rowptr[s]; aq <= bq; ++aq) hits[cur[col0 + cols[aq]]++] = (uint16_t)s;
        }
    out.tok.assign(nseg, {});
    out.lo.assign(nseg, {});
    out.mx.assign(nseg, 0);
    for (int t = 0; t < nseg; ++t) {
        const size_t e0 = (size_t)t * kSpSeg, e1 = std::min(S, e0 + kSpSeg);
        items.clear();
        for (size_t e = e0; e < e1; ++e) {
            const bool dg = diag_col[e] >= 0;
            const bool run = sparse_is_run(hits.data() + start[e], cnt[e]);
            const uint32_t cost = run ? 2 + (cnt[e] >> 4) : sparse_item_cost(cnt[e], !dg && cnt[e] <= 4);
            if (dg) {
                items.push_back({cost, (uint32_t)e, false});
                items.push_back({cost, (uint32_t)e, true});
            } else if (cnt[e] > 0) {
                items.push_back({cost, (uint32_t)e, false});
            }
        }
        std::sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
            if (a.cost != b.cost)
                return a.cost > b.cost;
            if (a.e != b.e)
                return a.e < b.e;
            return a.grad < b.grad;
        });
        // heap over (cost so far, rank); rank r is lane (r % kW) * 64 + r / kW
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> pq;
        for (int r = 0; r < kSpLanes; ++r) {
            pq.push({0, r});
            lane_items[r].clear();
        }
        for (const Item& it : items) {
            Load top = pq.top();
            pq.pop();
            lane_items[(top.second % kW) * 64 + top.second / kW].push_back(it);
            top.first += it.cost;
            pq.push(top);
        }
        uint64_t mx = 0;
        auto& tok = out.tok[t];
        auto& lo = out.lo[t];
        lo.assign(kSpLanes + 1, 0);
        for (int l = 0; l < kSpLanes; ++l) {
            lo[l] = (uint32_t)tok.size();
            uint64_t lc = 0;
            for (const Item& it : lane_items[l]) {
                const uint32_t e = it.e, off = (uint32_t)(e - e0), k = cnt[e];
                lc += it.cost;
                if (sparse_is_run(hits.data() + start[e], k)) {
                    const uint32_t kind = it.grad ? 2u : (diag_col[e] >= 0 ? 1u : 0u);
                    tok.push_back((uint16_t)off);  // code 0
                    tok.push_back((uint16_t)kind);
                    tok.push_back((uint16_t)k);
                    tok.push_back(hits[start[e]]);
                    if (it.grad)
                        tok.push_back((uint16_t)diag_col[e]);
                    continue;
                }
                if (it.grad) {
                    tok.push_back((uint16_t)(off | (15u << 12)));
                    tok.push_back((uint16_t)k);
                    tok.push_back((uint16_t)diag_col[e]);
                } else if (diag_col[e] >= 0) {
                    tok.push_back((uint16_t)(off | (14u << 12)));
                    tok.push_back((uint16_t)k);
                } else if (k <= 12) {
                    tok.push_back((uint16_t)(off | (k << 12)));
                } else {
                    tok.push_back((uint16_t)(off | (13u << 12)));
                    tok.push_back((uint16_t)k);
                }
                tok.insert(tok.end(), hits.begin() + start[e], hits.begin() + start[e] + k);
            }
            mx = std::max(mx, lc);
        }
        lo[kSpLanes] = (uint32_t)tok.size();
        out.mx[t] = mx;
    }
}

// The clients' parts are built on host threads (each client on its own, so
// the plan does not depend on the thread count), then laid out in client
// order; a client's segments are dispatched in decreasing order of their
// costliest lane (ties: lower segment).
static void build_sparse_plan(const double* features, const double* labels, int n, int m, int d, int m_pad,
                              SparsePlan& P)
{
    const size_t S = (size_t)d * (d + 1) / 2;
    const int nseg = (int)((S + kSpSeg - 1) / kSpSeg);
    std::vector<int> diag_col(S, -1);
    for (int j = 0; j < d; ++j) diag_col[(size_t)j * (j + 3) / 2] = j;
    std::vector<SparseClient> cl(n);
    {
        // clients taken from a counter by the calling thread and up to 15
        // more; a thread that cannot be started is simply not there. A
        // worker's failure (host memory) is raised here after the join.
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int nt = (int)std::min<unsigned>(std::min<unsigned>(hw, 16u), (unsigned)n);
        std::atomic<int> next{0};
        std::atomic<bool> failed{false};
        auto work = [&]() {
            try {
                for (int c; !failed.load() && (c = next.fetch_add(1)) < n;)
                    build_sparse_client(features, c, m, d, diag_col, cl[c]);
            } catch (...) {
                failed.store(true);
            }
        };
        std::vector<std::thread> pool;
        try {
            for (int t = 1; t < nt; ++t) pool.emplace_back(work);
        } catch (...) {
        }
        work();
        for (auto& th : pool) th.join();
        if (failed.load())
            throw std::bad_alloc();
    }
    P.seg.assign((size_t)n * nseg * 4, 0);
    P.lane.assign((size_t)n * nseg * (kSpLanes + 1), 0);
    P.stream.clear();
    P.b.assign((size_t)n * m_pad, 0.0);
    P.rowptr.assign((size_t)n * (m + 1), 0);
    P.cols.clear();
    P.max_tokens = 0;
    std::vector<int> order(nseg);
    for (int c = 0; c < n; ++c) {
        SparseClient& sc = cl[c];
        for (int s = 0; s < m; ++s) P.b[(size_t)c * m_pad + s] = labels[(size_t)c * m + s] == 1.0 ? 1.0 : -1.0;
        const uint32_t cbase = (uint32_t)P.cols.size();
        for (int s = 0; s <= m; ++s) P.rowptr[(size_t)c * (m + 1) + s] = cbase + sc.rowptr[s];
        P.cols.insert(P.cols.end(), sc.cols.begin(), sc.cols.end());
        for (int t = 0; t < nseg; ++t) order[t] = t;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sc.mx[a] > sc.mx[b]; });
        for (int r = 0; r < nseg; ++r) {
            const int t = order[r];
            while (P.stream.size() % 4) P.stream.push_back(0);
            const uint64_t off = P.stream.size();
            uint32_t* rec = &P.seg[((size_t)c * nseg + r) * 4];
            rec[0] = (uint32_t)t;
            rec[1] = (uint32_t)sc.tok[t].size();
            rec[2] = (uint32_t)(off & 0xFFFFFFFFull);
            rec[3] = (uint32_t)(off >> 32);
            std::copy(sc.lo[t].begin(), sc.lo[t].end(), P.lane.begin() + ((size_t)c * nseg + r) * (kSpLanes + 1));
            P.stream.insert(P.stream.end(), sc.tok[t].begin(), sc.tok[t].end());
            P.max_tokens = std::max<size_t>(P.max_tokens, sc.tok[t].size());
        }
        sc = SparseClient();  // release as we go
    }
    P.stream.resize(P.stream.size() + 8, 0);  // the kernel copies whole 8-byte words
    P.cols.push_back(0);                        // never empty
}

// LDS the sparse form may take per workgroup
constexpr size_t kSpMaxLds = 160 * 1024 - 1024;

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
    if (d > kMaxD) {
        set_err("fednl_create: d > 2048");
        return FEDNL_E_UNSUPPORTED;
    }
    if (opts->hessian_mode < 0 || opts->hessian_mode > 2) {
        set_err("fednl_create: hessian_mode must be 0 (FP64), 1 (FP32) or 2 (BF16 split)");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (opts->hessian_form < FEDNL_HESS_AUTO || opts->hessian_form > FEDNL_HESS_SPARSE) {
        set_err("fednl_create: hessian_form must be FEDNL_HESS_AUTO, FEDNL_HESS_DENSE or FEDNL_HESS_SPARSE");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (opts->line_search &&
        (!(opts->c_line_search > 0.0) || !(opts->gamma_line_search > 0.0 && opts->gamma_line_search < 1.0) ||
         opts->max_line_search_steps < 0 || opts->max_line_search_steps > kMaxLineSearchSteps)) {
        set_err("fednl_create: line search needs c > 0, 0 < gamma < 1 and 0 <= max steps <= 64");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (opts->clients_per_round < 0 || opts->clients_per_round > n) {
        set_err("fednl_create: clients_per_round must be in [1, n] (FedNL-PP) or 0 (FedNL)");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (opts->algorithm != FEDNL_ALG_FEDNL && opts->algorithm != FEDNL_ALG_GD) {
        set_err("fednl_create: algorithm must be FEDNL_ALG_FEDNL or FEDNL_ALG_GD");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (opts->algorithm == FEDNL_ALG_GD && (opts->line_search || opts->clients_per_round > 0)) {
        // the reference parses the line-search flags only for FedNL (Args.h:321-380)
        set_err("fednl_create: GD runs without line search and without FedNL-PP");
        return FEDNL_E_UNSUPPORTED;
    }
    if (opts->clients_per_round > 0 && opts->line_search) {
        // bin_fednl_local_pp has no line search (its Args.h drops the flags)
        set_err("fednl_create: FedNL-PP runs without line search");
        return FEDNL_E_UNSUPPORTED;
    }
    Engine* e = new Engine();
    e->hmode = opts->hessian_mode;
    e->o = *opts;
    const bool gd = opts->algorithm == FEDNL_ALG_GD;
    if (gd) {
        // GD sends gradients only: no Hessians, no selection (the smallest
        // selection buffers, never launched)
        e->hmode = 0;
        e->o.hessian_mode = 0;
        e->o.compressor = FEDNL_RANDSEQK;
    }
    e->n = n;
    e->m = m;
    e->d = d;
    e->lda = round_up(d, 4);
    e->S = (size_t)d * (d + 1) / 2;
    const int comp = e->o.compressor;
    if (gd)
        e->k = 1;
    else if (comp == FEDNL_IDENTICAL || comp == FEDNL_NATURAL)
        e->k = (int)e->S;  // TrainLoop.h:354-377
    else
        e->k = (int)(opts->k_mult * d);  // TrainLoop.h:362
    if (e->k <= 0 || (size_t)e->k > e->S) {
        set_err("fednl_create: k must be in [1, d(d+1)/2]");
        delete e;
        return FEDNL_E_WRONG_ARGUMENT;
    }
    if (comp != FEDNL_TOPK && comp != FEDNL_TOPLEK && comp != FEDNL_RANDK && comp != FEDNL_RANDSEQK &&
        comp != FEDNL_IDENTICAL && comp != FEDNL_NATURAL) {
        set_err("fednl_create: compressor not supported by this engine");
        delete e;
        return FEDNL_E_UNSUPPORTED;
    }
    if (select_keeps_k(comp, e->k, e->S) && e->k > kMaxSelK) {
        set_err("fednl_create: k above the selection limit (32768)");
        delete e;
        return FEDNL_E_UNSUPPORTED;
    }
    e->m_pad = round_up(m, kKB);  // a multiple of every kernel's sample chunk
    e->d_pad = round_up(d, kTile);
    e->nt = e->d_pad / kTile;
    e->tiles_up = e->nt * (e->nt + 1) / 2;
    e->nparts = e->tiles_up;
    e->dp = round_up(d, kNB);
    e->workers = std::max(1, std::min(opts->workers, n));
    e->alpha = theory_alpha(comp, opts->alpha_mode, opts->alpha, e->k, d, e->S);
    e->inv_n = 1.0 / (double)n;
    e->prescale = e->inv_n * e->alpha;
    e->mult = (double)e->S / (double)e->k;
    e->lek_delta = (double)e->k * (1.0 / (double)e->S);
    e->wide = select_wide(comp, e->k, e->S);
    e->sel_scr_stride = e->wide ? select_scratch_stride(comp, e->k, e->S) : 0;
    e->lds_select = select_lds(comp, e->k, e->S, e->wide, big_ring(e, false));
    e->lds_select_b = select_lds(comp, e->k, e->S, e->wide, big_ring(e, true));
    e->lds_emit = emit_lds(e->S, e->wide);
    e->lds_newton = mc_lds_bytes(e->dp);

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
    {
        const char* ge = getenv("FEDNL_ROUND_GRAPH");
        if (ge && ge[0] == '1')
            e->use_graph = true;
    }
    {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
            e->wclk_hz = 1e3 * khz;
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
    ALLOC(e->part, (size_t)n * std::max<size_t>(e->tiles_up, (e->S + kSpSeg - 1) / kSpSeg) * 2);
    ALLOC(e->lk, n);
    ALLOC(e->sel, (size_t)n * e->k);
    ALLOC(e->pos, (size_t)n * e->k);
    ALLOC(e->idx, (size_t)n * e->k);
    ALLOC(e->vals, (size_t)n * e->k);
    ALLOC(e->rng, n);
    ALLOC(e->lek_state, e->workers);
    ALLOC(e->lek_popped, n);
    ALLOC(e->seq_start, n);
    e->bm_words = (int)((e->S + 31) / 32);
    ALLOC(e->sel_bm, (size_t)n * e->bm_words);
    if (e->sel_scr_stride)
        ALLOC(e->sel_scr, (size_t)n * e->sel_scr_stride);
    ALLOC(e->sel_cnt, n);
    ALLOC(e->cnt, n);
    ALLOC(e->lek_hit, n);
    ALLOC(e->lek_p, n);
    ALLOC(e->round_ctr, 1);
    ALLOC(e->H, e->S);
    ALLOC(e->g_avg, e->d_pad);
    ALLOC(e->shift, 1);
    ALLOC(e->W, (size_t)e->dp * e->dp);
    ALLOC(e->nsync, (size_t)mc_sync_words(e->dp));
    ALLOC(e->ny, (size_t)e->dp);
    ALLOC(e->ninv, (size_t)e->dp);
    ALLOC(e->nxt, (size_t)e->dp * kNB);
    if (hipMemset(e->nsync, 0, sizeof(int) * mc_sync_words(e->dp)) != hipSuccess) {
        set_err("hipMemset Newton sync words");
        free_engine(e);
        return FEDNL_E_DEVICE;
    }
    {
        // W padding: identity (rows/columns d..dp-1), zero elsewhere; the lower
        // d x d part is rewritten every round
        std::vector<double> w0((size_t)e->dp * e->dp, 0.0);
        for (int i = d; i < e->dp; ++i) w0[(size_t)i * e->dp + i] = 1.0;
        if (hipMemcpy(e->W, w0.data(), sizeof(double) * w0.size(), hipMemcpyHostToDevice) != hipSuccess) {
            set_err("hipMemcpy W");
            free_engine(e);
            return FEDNL_E_DEVICE;
        }
    }
    ALLOC(e->dx, e->d_pad);
    ALLOC(e->mst, kLsHead + (size_t)e->d_pad);
    ALLOC(e->H0, e->S);
    if (e->hmode != 0 && opts->clients_per_round <= 0)
        ALLOC(e->gpart, (size_t)n * (e->m_pad / 16) * e->d_pad);
    if (opts->clients_per_round > 0) {
        ALLOC(e->pp_sel, n);
        ALLOC(e->pp_rng, 1);
        ALLOC(e->pp_dir, (size_t)n * e->d_pad);
        ALLOC(e->pp_lk, n);
        ALLOC(e->pp_gmsg, (size_t)n * e->d_pad);
        ALLOC(e->pp_lkmsg, n);
        ALLOC(e->pp_tg, e->d_pad);
        if (opts->option_b)
            ALLOC(e->pp_h, (size_t)n * e->S);
    }
#undef ALLOC
    if ((rc = ensure_trace(e, 1024)) != FEDNL_OK) {
        free_engine(e);
        return rc;
    }
    {
        const size_t lds = std::max(e->lds_select, e->lds_emit);
        hipError_t le = hipSuccess;
        const void* sel_fns[] = {
            (const void*)k_select_heap<false, true, false>, (const void*)k_select_heap<false, false, false>,
            (const void*)k_select_heap<true, false, false>, (const void*)k_select_randk<true, false>,
            (const void*)k_select_randk<false, false>,      (const void*)k_select_window<true, false>,
            (const void*)k_emit<false>,                     (const void*)k_select_heap<false, true, true>,
            (const void*)k_select_heap<false, false, true>, (const void*)k_select_heap<true, false, true>};
        for (const void* f : sel_fns)
            if (le == hipSuccess)
                le = raise_lds_limit(f, lds);
        if (le == hipSuccess)
            le = raise_lds_limit((const void*)k_newton_mc, e->lds_newton);
        if (le == hipSuccess && e->gpart)
            le = raise_lds_limit((const void*)k_margin<false, 4, true>, margin_grad_lds(e->d_pad));
        if (le != hipSuccess) {
            set_err("hipFuncSetAttribute (LDS limit)", le);
            free_engine(e);
            return FEDNL_E_DEVICE;
        }
    }

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
    // the FP64 Hessian's form: sparse for binary designs with few
    // co-occurring feature pairs (at most a tenth of the dense m S terms)
    if (!gd) {
        int want = opts->hessian_form;
        const char* fe = getenv("FEDNL_HESS_FORM");
        if (fe && !strcmp(fe, "dense"))
            want = FEDNL_HESS_DENSE;
        else if (fe && !strcmp(fe, "sparse"))
            want = FEDNL_HESS_SPARSE;
        const size_t cells = (size_t)n * m * d;
        const bool binary = e->hmode == 0 && m <= 65535 && want != FEDNL_HESS_DENSE && design_is_binary(features, cells);
        bool sparse = false;
        if (binary) {
            const double pairs = sparse_pair_count(features, (size_t)n * m, d);
            sparse = want == FEDNL_HESS_SPARSE || pairs <= 0.1 * (double)n * (double)m * (double)e->S;
        }
        if (want == FEDNL_HESS_SPARSE && !sparse) {
            set_err("fednl_create: the sparse Hessian form needs FP64 mode, m <= 65535 and 0/1 features");
            free_engine(e);
            return FEDNL_E_UNSUPPORTED;
        }
        if (sparse) {
            SparsePlan plan;
            try {
                build_sparse_plan(features, labels, n, m, d, e->m_pad, plan);
            } catch (...) {
                set_err("fednl_create: out of host memory building the sparse Hessian form's sample lists");
                free_engine(e);
                return FEDNL_E_INTERNAL;
            }
            const size_t lds = sparse_lds_bytes(e->m_pad, e->d_pad, plan.max_tokens);
            if (lds > kSpMaxLds) {
                if (want == FEDNL_HESS_SPARSE) {
                    set_err("fednl_create: a segment's sample lists exceed the LDS of the sparse Hessian form");
                    free_engine(e);
                    return FEDNL_E_UNSUPPORTED;
                }
                sparse = false;
            } else {
                hipError_t se = hipMalloc((void**)&e->sp_seg, sizeof(uint32_t) * plan.seg.size());
                if (se == hipSuccess)
                    se = hipMalloc((void**)&e->sp_lane, sizeof(uint32_t) * plan.lane.size());
                if (se == hipSuccess)
                    se = hipMalloc((void**)&e->sp_stream, sizeof(uint16_t) * plan.stream.size());
                if (se == hipSuccess)
                    se = hipMalloc((void**)&e->sp_b, sizeof(double) * plan.b.size());
                if (se == hipSuccess)
                    se = hipMemcpy(e->sp_seg, plan.seg.data(), sizeof(uint32_t) * plan.seg.size(),
                                   hipMemcpyHostToDevice);
                if (se == hipSuccess)
                    se = hipMemcpy(e->sp_lane, plan.lane.data(), sizeof(uint32_t) * plan.lane.size(),
                                   hipMemcpyHostToDevice);
                if (se == hipSuccess)
                    se = hipMemcpy(e->sp_stream, plan.stream.data(), sizeof(uint16_t) * plan.stream.size(),
                                   hipMemcpyHostToDevice);
                if (se == hipSuccess)
                    se = hipMemcpy(e->sp_b, plan.b.data(), sizeof(double) * plan.b.size(), hipMemcpyHostToDevice);
                if (se == hipSuccess)
                    se = hipMalloc((void**)&e->sp_rowptr, sizeof(uint32_t) * plan.rowptr.size());
                if (se == hipSuccess)
                    se = hipMalloc((void**)&e->sp_cols, sizeof(uint16_t) * plan.cols.size());
                if (se == hipSuccess)
                    se = hipMemcpy(e->sp_rowptr, plan.rowptr.data(), sizeof(uint32_t) * plan.rowptr.size(),
                                   hipMemcpyHostToDevice);
                if (se == hipSuccess)
                    se = hipMemcpy(e->sp_cols, plan.cols.data(), sizeof(uint16_t) * plan.cols.size(),
                                   hipMemcpyHostToDevice);
                if (se == hipSuccess)
                    se = raise_lds_limit((const void*)k_hessian_sparse<true>, lds);
                if (se == hipSuccess)
                    se = raise_lds_limit((const void*)k_hessian_sparse<false>, lds);
                if (se != hipSuccess) {
                    set_err("sparse Hessian form upload", se);
                    free_engine(e);
                    return FEDNL_E_DEVICE;
                }
                e->lds_sparse = lds;
                e->sp_bytes = sizeof(uint32_t) * (plan.seg.size() + plan.lane.size() + plan.rowptr.size()) +
                              sizeof(uint16_t) * (plan.stream.size() + plan.cols.size()) +
                              sizeof(double) * plan.b.size();
                e->hform = FEDNL_HESS_SPARSE;
                e->nparts = (int)((e->S + kSpSeg - 1) / kSpSeg);
            }
        }
    }
    // reduced-precision copies of A for the Hessian contraction
    if (e->hmode != 0) {
        const size_t nA = (size_t)n * e->m_pad * e->d_pad;
        hipError_t ae = hipSuccess;
        if (e->hmode == 1)
            ae = hipMalloc((void**)&e->A32, sizeof(float) * nA);
        else {
            ae = hipMalloc((void**)&e->AThi, sizeof(__bf16) * nA);
            if (ae == hipSuccess)
                ae = hipMalloc((void**)&e->ATlo, sizeof(__bf16) * nA);
        }
        if (ae != hipSuccess) {
            set_err("hipMalloc (reduced-precision A)", ae);
            free_engine(e);
            return FEDNL_E_DEVICE;
        }
        if (e->hmode == 1)
            hipLaunchKernelGGL(k_convert_a, dim3(2048), dim3(256), 0, e->st, e->A, nA, e->A32);
        else
            hipLaunchKernelGGL(k_convert_bf16t, dim3((unsigned)((size_t)n * (e->m_pad / 32) * (e->d_pad / 64))),
                               dim3(256), 0, e->st, e->A, e->m_pad, e->d_pad, e->AThi, e->ATlo);
    }
    std::vector<uint32_t> seeds(n);
    for (int c = 0; c < n; ++c) seeds[c] = (uint32_t)(123 + c);  // TrainLoop.h:632
    std::vector<uint32_t> lek(e->workers, 1234u);                   // RandomGenRealLinear default seed
    hipMemcpyAsync(e->rng, seeds.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, e->st);
    hipMemcpyAsync(e->lek_state, lek.data(), sizeof(uint32_t) * e->workers, hipMemcpyHostToDevice, e->st);
    hipMemsetAsync(e->x, 0, sizeof(double) * e->d_pad, e->st);
    hipMemsetAsync(e->round_ctr, 0, sizeof(int), e->st);
    hipMemsetAsync(e->mst, 0, sizeof(double) * kLsHead, e->st);
    hipMemsetAsync(e->lek_hit, 0, sizeof(int) * n, e->st);
    hipMemsetAsync(e->grad_c, 0, sizeof(double) * (size_t)n * e->d_pad, e->st);

    // learned_c = H_c(x0), H = mean_c learned_c
    if (hipMalloc((void**)&e->pdev, sizeof(ProblemDev)) != hipSuccess || sync_desc(e) != FEDNL_OK) {
        set_err("descriptor upload (create)");
        free_engine(e);
        return FEDNL_E_DEVICE;
    }
    {
        tune_margin(e);
        if (!gd) {
            launch_margin(e, false);
            launch_hessian(e, true, nullptr);
            hipLaunchKernelGGL(k_init_H, dim3((unsigned)((e->S + 255) / 256)), dim3(256), 0, e->st, e->H,
                               e->learned, e->S, n, e->inv_n);
            hipMemcpyAsync(e->H0, e->H, sizeof(double) * e->S, hipMemcpyDeviceToDevice, e->st);
            tune_hessian(e);
        }
        if (e->pp_sel) {
            // FedNL-PP: every client's gradient and direction at x0, their
            // average, l_c = 0, the client sampler's seed (Args.h:242-244)
            const ProblemDev* pd = e->pdev;
            hipMemsetAsync(e->pp_sel, 0, sizeof(int) * n, e->st);
            hipMemsetAsync(e->pp_lk, 0, sizeof(double) * n, e->st);
            hipMemsetAsync(e->pp_lkmsg, 0, sizeof(double) * n, e->st);
            hipMemsetAsync(e->pp_gmsg, 0, sizeof(double) * (size_t)n * e->d_pad, e->st);
            hipMemsetAsync(e->pp_dir, 0, sizeof(double) * (size_t)n * e->d_pad, e->st);
            hipMemsetAsync(e->pp_tg, 0, sizeof(double) * e->d_pad, e->st);
            const uint32_t seed = opts->client_selection_seed;
            hipMemcpyAsync(e->pp_rng, &seed, sizeof(uint32_t), hipMemcpyHostToDevice, e->st);
            hipLaunchKernelGGL(k_grad_f64, dim3(blocks_grad(e->hdesc)), dim3(256), 0, e->st, pd, (const int2*)nullptr);
            hipLaunchKernelGGL(k_pp_client<true>, dim3(n), dim3(256), 0, e->st, pd, (const int2*)nullptr);
            hipLaunchKernelGGL(k_pp_init_avg, dim3(1), dim3(256), 0, e->st, pd, (const int2*)nullptr);
            hipStreamSynchronize(e->st);  // the seed's source is on this stack frame
        }
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
    if ((rc = sync_desc(e)) != FEDNL_OK)
        return rc;
    const bool graph = e->use_graph && !e->timing;
    if (graph && !e->gexec) {
        // every kernel of a round reads the problem through e->pdev, whose
        // address and the launch shapes never change: one capture serves
        // all rounds
        HIP_OK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
        launch_round(e);
        hipError_t ce = hipStreamEndCapture(e->st, &e->graph);
        if (ce == hipSuccess)
            ce = hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            e->use_graph = false;
            if (e->graph) {
                hipGraphDestroy(e->graph);
                e->graph = nullptr;
            }
            e->gexec = nullptr;
            hipGetLastError();
        }
    }
    for (int r = 0; r < rounds; ++r) {
        if (graph && e->gexec)
            HIP_OK(hipGraphLaunch(e->gexec, e->st));
        else
            launch_round(e);
        e->rounds_enqueued += 1;
    }
    HIP_OK(hipGetLastError());
    return FEDNL_OK;
}

// ---------------------------------------------------------------------------
// Batch: many independent problems (engines) stepped together. Every kernel
// of a round is launched once for all members, over a block map that names
// (member, local block) for each workgroup, so small problems share the chip:
// e.g. the members' Newton solves run as concurrent workgroups. A round can be
// captured once into a HIP graph and replayed.
// ---------------------------------------------------------------------------
enum BatchKernel {
    BK_MARGIN, BK_MARGIN_G, BK_MARGIN_SP, BK_HESS64, BK_HESS_SP, BK_HESS32, BK_HESS16, BK_GRAD, BK_SEL_TOPK, BK_SEL_TOPLEK, BK_SEL_RANDK, BK_SEL_WINDOW,
    BK_EMIT, BK_SEL_TOPK_W, BK_SEL_TOPLEK_W, BK_SEL_RANDK_W, BK_SEL_WINDOW_W, BK_EMIT_W, BK_MASTER, BK_NEWTON,
    BK_LS_MARGIN, BK_LS_MARGIN_SP, BK_LS, BK_COUNT
};

struct Batch {
    std::vector<Engine*> m;
    hipStream_t st = nullptr;
    ProblemDev* d_probs = nullptr;
    std::vector<ProblemDev> h_probs;
    int2* d_maps = nullptr;
    int off[BK_COUNT] = {0}, cnt[BK_COUNT] = {0};
    size_t lds[BK_COUNT] = {0};  // dynamic LDS of each launch: the most any member needs
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int rounds = 0;
    int ls_steps = 0;  // most line-search trials of any member
};

static void free_batch(Batch* b)
{
    if (!b)
        return;
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->d_probs) hipFree(b->d_probs);
    if (b->d_maps) hipFree(b->d_maps);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

static int batch_build(Batch* b)
{
    std::vector<int2> maps[BK_COUNT];
    for (int q = 0; q < BK_COUNT; ++q) b->lds[q] = 0;
    auto add = [&](int kind, int p, int nb, bool align8, size_t lds = 0) {
        b->lds[kind] = std::max(b->lds[kind], lds);
        auto& v = maps[kind];
        if (align8)
            while (v.size() % 8) v.push_back(make_int2(-1, 0));
        for (int q = 0; q < nb; ++q) v.push_back(make_int2(p, q));
    };
    b->h_probs.clear();
    for (size_t i = 0; i < b->m.size(); ++i) {
        Engine* e = b->m[i];
        const ProblemDev P = make_desc(e, true);
        b->h_probs.push_back(P);
        const int p = (int)i;
        if (e->hform == FEDNL_HESS_SPARSE)
            add(BK_MARGIN_SP, p, blocks_margin_sparse(P), false, sizeof(double) * (size_t)e->d_pad);
        else if (e->gpart)
            add(BK_MARGIN_G, p, blocks_margin(P), false, margin_grad_lds(e->d_pad));
        else
            add(BK_MARGIN, p, blocks_margin(P), false, sizeof(double) * (size_t)e->d_pad);
        if (e->hform == FEDNL_HESS_SPARSE)
            add(BK_HESS_SP, p, blocks_sparse(P), true, e->lds_sparse);
        else
            add(e->hmode == 0 ? BK_HESS64 : (e->hmode == 1 ? BK_HESS32 : BK_HESS16), p, blocks_hessian(P), true);
        if (e->hmode != 0 && !e->gpart)
            add(BK_GRAD, p, blocks_grad(P), false);
        const int comp = e->o.compressor;
        const bool w = e->wide;
        // selection (with the messages fused, except for TopLEK)
        const size_t fused = std::max(e->lds_select_b, e->lds_emit);
        if (comp == FEDNL_TOPK)
            add(w ? BK_SEL_TOPK_W : BK_SEL_TOPK, p, e->n, false, fused);
        else if (comp == FEDNL_TOPLEK)
            add(w ? BK_SEL_TOPLEK_W : BK_SEL_TOPLEK, p, e->n, false, e->lds_select_b);
        else if (comp == FEDNL_RANDK)
            add(w ? BK_SEL_RANDK_W : BK_SEL_RANDK, p, e->n, false, fused);
        else
            add(w ? BK_SEL_WINDOW_W : BK_SEL_WINDOW, p, e->n, false, e->lds_emit);
        if (comp == FEDNL_TOPLEK)
            add(w ? BK_EMIT_W : BK_EMIT, p, e->n, false, e->lds_emit);
        add(BK_MASTER, p, blocks_master(P), false);
        add(BK_NEWTON, p, mc_blocks(P.dp), false, e->lds_newton);
        if (e->o.line_search) {
            if (e->hform == FEDNL_HESS_SPARSE)
                add(BK_LS_MARGIN_SP, p, blocks_margin_sparse(P), false, sizeof(double) * (size_t)e->d_pad);
            else
                add(BK_LS_MARGIN, p, blocks_margin(P), false, sizeof(double) * (size_t)e->d_pad);
            add(BK_LS, p, 1, false);
            b->ls_steps = std::max(b->ls_steps, e->o.max_line_search_steps);
        }
    }
    std::vector<int2> all;
    for (int k = 0; k < BK_COUNT; ++k) {
        // every segment starts at a multiple of 8 entries
        while (all.size() % 8) all.push_back(make_int2(-1, 0));
        b->off[k] = (int)all.size();
        b->cnt[k] = (int)maps[k].size();
        all.insert(all.end(), maps[k].begin(), maps[k].end());
    }
    HIP_OK(hipMalloc((void**)&b->d_probs, sizeof(ProblemDev) * std::max<size_t>(1, b->h_probs.size())));
    HIP_OK(hipMalloc((void**)&b->d_maps, sizeof(int2) * std::max<size_t>(1, all.size())));
    HIP_OK(hipMemcpy(b->d_probs, b->h_probs.data(), sizeof(ProblemDev) * b->h_probs.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->d_maps, all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice));
    return FEDNL_OK;
}

// the members' descriptors again (trace buffers may have grown)
static int batch_sync_desc(Batch* b)
{
    bool changed = false;
    for (size_t i = 0; i < b->m.size(); ++i) {
        const ProblemDev P = make_desc(b->m[i], true);
        if (memcmp(&P, &b->h_probs[i], sizeof(P)) != 0) {
            b->h_probs[i] = P;
            changed = true;
        }
    }
    if (changed) {
        HIP_OK(hipMemcpyAsync(b->d_probs, b->h_probs.data(), sizeof(ProblemDev) * b->h_probs.size(),
                              hipMemcpyHostToDevice, b->st));
        HIP_OK(hipStreamSynchronize(b->st));
    }
    return FEDNL_OK;
}

static void batch_launch_round(Batch* b)
{
    const ProblemDev* pd = b->d_probs;
    const int2* mp = b->d_maps;
    hipStream_t st = b->st;
    auto map = [&](int k) { return mp + b->off[k]; };
    if (b->cnt[BK_MARGIN])
        hipLaunchKernelGGL((k_margin<false, 4>), dim3(b->cnt[BK_MARGIN]), dim3(256), b->lds[BK_MARGIN], st, pd,
                           map(BK_MARGIN));
    if (b->cnt[BK_MARGIN_SP])
        hipLaunchKernelGGL(k_margin_sparse<false>, dim3(b->cnt[BK_MARGIN_SP]), dim3(kSpMarginRows),
                           b->lds[BK_MARGIN_SP], st, pd, map(BK_MARGIN_SP));
    if (b->cnt[BK_MARGIN_G])
        hipLaunchKernelGGL((k_margin<false, 4, true>), dim3(b->cnt[BK_MARGIN_G]), dim3(256), b->lds[BK_MARGIN_G], st,
                           pd, map(BK_MARGIN_G));
    if (b->cnt[BK_HESS64])
        hipLaunchKernelGGL((k_hessian_f64<false, 16>), dim3(b->cnt[BK_HESS64]), dim3(256), 0, st, pd, map(BK_HESS64),
                           (double*)nullptr);
    if (b->cnt[BK_HESS_SP])
        hipLaunchKernelGGL(k_hessian_sparse<false>, dim3(b->cnt[BK_HESS_SP]), dim3(kSpThreads), b->lds[BK_HESS_SP], st, pd,
                           map(BK_HESS_SP), (double*)nullptr);
    if (b->cnt[BK_GRAD])
        hipLaunchKernelGGL(k_grad_f64, dim3(b->cnt[BK_GRAD]), dim3(256), 0, st, pd, map(BK_GRAD));
    if (b->cnt[BK_HESS32])
        hipLaunchKernelGGL(k_hessian_f32<false>, dim3(b->cnt[BK_HESS32]), dim3(256), 0, st, pd, map(BK_HESS32),
                           (double*)nullptr);
    if (b->cnt[BK_HESS16])
        hipLaunchKernelGGL(k_hessian_bf16<false>, dim3(b->cnt[BK_HESS16]), dim3(256), 0, st, pd, map(BK_HESS16),
                           (double*)nullptr);
    // selection and messages in one workgroup per client (TopLEK: emit after)
#define BK_LAUNCH(kind, ...)                                                                         \
    if (b->cnt[kind])                                                                                \
        hipLaunchKernelGGL((__VA_ARGS__), dim3(b->cnt[kind]), dim3(256), b->lds[kind], st, pd, map(kind))
    BK_LAUNCH(BK_SEL_TOPK, k_select_heap<false, true, false>);
    BK_LAUNCH(BK_SEL_TOPLEK, k_select_heap<true, false, false>);
    BK_LAUNCH(BK_SEL_RANDK, k_select_randk<true, false>);
    BK_LAUNCH(BK_SEL_WINDOW, k_select_window<true, false>);
    BK_LAUNCH(BK_EMIT, k_emit<false>);
    BK_LAUNCH(BK_SEL_TOPK_W, k_select_heap<false, true, true>);
    BK_LAUNCH(BK_SEL_TOPLEK_W, k_select_heap<true, false, true>);
    BK_LAUNCH(BK_SEL_RANDK_W, k_select_randk<true, true>);
    BK_LAUNCH(BK_SEL_WINDOW_W, k_select_window<true, true>);
    BK_LAUNCH(BK_EMIT_W, k_emit<true>);
#undef BK_LAUNCH
    hipLaunchKernelGGL(k_master, dim3(b->cnt[BK_MASTER]), dim3(256), 0, st, pd, map(BK_MASTER));
    hipLaunchKernelGGL(k_newton_mc, dim3(b->cnt[BK_NEWTON]), dim3(kMcThreads), b->lds[BK_NEWTON], st, pd,
                       map(BK_NEWTON));
    if (b->cnt[BK_LS]) {
        hipLaunchKernelGGL(k_ls_begin, dim3(b->cnt[BK_LS]), dim3(256), 0, st, pd, map(BK_LS));
        for (int t = 0; t < b->ls_steps; ++t) {
            if (b->cnt[BK_LS_MARGIN])
                hipLaunchKernelGGL((k_margin<true, 4>), dim3(b->cnt[BK_LS_MARGIN]), dim3(256), b->lds[BK_LS_MARGIN],
                                   st, pd, map(BK_LS_MARGIN));
            if (b->cnt[BK_LS_MARGIN_SP])
                hipLaunchKernelGGL(k_margin_sparse<true>, dim3(b->cnt[BK_LS_MARGIN_SP]), dim3(kSpMarginRows),
                                   b->lds[BK_LS_MARGIN_SP], st, pd, map(BK_LS_MARGIN_SP));
            hipLaunchKernelGGL(k_ls_check, dim3(b->cnt[BK_LS]), dim3(256), 0, st, pd, map(BK_LS));
        }
    }
}

static int batch_enqueue(Batch* b, int rounds)
{
    if (rounds <= 0 || b->m.empty())
        return FEDNL_OK;
    // a trace buffer that grows is copied on the member's stream: the batch
    // stream must not be writing the old one meanwhile
    for (Engine* e : b->m)
        if (e->rounds_enqueued + rounds + 1 > e->trace_cap) {
            HIP_OK(hipStreamSynchronize(b->st));
            break;
        }
    for (Engine* e : b->m) {
        int rc = ensure_trace(e, e->rounds_enqueued + rounds + 1);
        if (rc != FEDNL_OK)
            return rc;
    }
    int rc = batch_sync_desc(b);
    if (rc != FEDNL_OK)
        return rc;
    if (b->use_graph && !b->gexec) {
        HIP_OK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        batch_launch_round(b);
        hipError_t ce = hipStreamEndCapture(b->st, &b->graph);
        if (ce == hipSuccess)
            ce = hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            // fall back to plain launches on this stream
            b->use_graph = false;
            if (b->graph) {
                hipGraphDestroy(b->graph);
                b->graph = nullptr;
            }
            b->gexec = nullptr;
            hipGetLastError();
        }
    }
    for (int r = 0; r < rounds; ++r) {
        if (b->gexec)
            HIP_OK(hipGraphLaunch(b->gexec, b->st));
        else
            batch_launch_round(b);
        for (Engine* e : b->m)
            e->rounds_enqueued += 1;
        b->rounds += 1;
    }
    HIP_OK(hipGetLastError());
    return FEDNL_OK;
}

// device bytes one engine holds for a problem of this size (what the engine
// allocates; the admission test of the serving loop)
static size_t engine_footprint(const fednl_opts* o, int n, int m, int d)
{
    const size_t m_pad = (size_t)round_up(m, kKB), d_pad = (size_t)round_up(d, kTile), dp = (size_t)round_up(d, kNB);
    const size_t S = (size_t)d * (d + 1) / 2;
    size_t k = (o->compressor == FEDNL_IDENTICAL || o->compressor == FEDNL_NATURAL) ? S : (size_t)(o->k_mult * d);
    size_t b = 8 * (size_t)n * m_pad * d_pad;                 // A
    if (o->hessian_mode == 1) b += 4 * (size_t)n * m_pad * d_pad;
    if (o->hessian_mode == 2) b += 4 * (size_t)n * m_pad * d_pad;
    b += 8 * 2 * (size_t)n * S;                                 // learned, delta
    b += 8 * 3 * (size_t)n * m_pad + 8 * (size_t)n * d_pad;     // Dw, gw, nlog, grad_c
    b += (size_t)n * k * (4 + 4 + 4 + 8) + 4 * (size_t)n * ((S + 31) / 32);
    b += 8 * (2 * S + dp * dp) + (1 << 20);                     // H, H0, W
    if (o->clients_per_round > 0)                               // FedNL-PP state
        b += 8 * (2 * (size_t)n * d_pad + 2 * (size_t)n + d_pad) + (o->option_b ? 8 * (size_t)n * S : 0);
    return b;
}

// ---------------------------------------------------------------------------
// Newton solve without allocation. Every buffer of the standalone solve that
// is neither an input nor an output (the descriptor, the shift, the status
// record, the round counter and the Cholesky workspace W) lives in one
// file-scope device array. Each GPU holds its own copy until the process
// ends; its device address comes from hipGetSymbolAddress once per device
// (the array's name in host code is only a stand-in). Calls on one GPU share
// the array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int kSolveScratchMaxD = kMaxNewtonD;                 // d <= 512
constexpr size_t kSolveHdrDoubles = 256;                       // 2 KiB header
constexpr size_t kSolveSlotShift = 128, kSolveSlotTrace = 136, kSolveSlotCtr = 152, kSolveSlotSync = 160;
constexpr size_t kSolveOffNy = kSolveHdrDoubles;
constexpr size_t kSolveOffNinv = kSolveOffNy + kSolveScratchMaxD;
constexpr size_t kSolveOffNxt = kSolveOffNinv + kSolveScratchMaxD;
constexpr size_t kSolveOffW = kSolveOffNxt + (size_t)kSolveScratchMaxD * kNB;
constexpr size_t kSolveScratchDoubles =
    kSolveOffW + (size_t)kSolveScratchMaxD * (size_t)kSolveScratchMaxD;  // 2 MiB + 138 KiB
static_assert(sizeof(ProblemDev) <= kSolveSlotShift * sizeof(double), "descriptor fits the header");
static_assert(kSolveSlotTrace + kTraceStride <= kSolveSlotCtr, "status record fits before the counter");
static_assert(kSolveSlotSync * 2 + mc_sync_words(kSolveScratchMaxD) <= kSolveHdrDoubles * 2, "sync words fit");
static_assert(kSolveOffW % 32 == 0, "W starts on a 256-byte boundary");

static __device__ __align__(256) double scratch[kSolveScratchDoubles];

// One launch before the solve: the descriptor (by value) and the shift into
// the header, the status record, round counter and sync words cleared, and W
// built from the packed upper H (lower triangle, column-major, identity
// padding: what the allocating solve uploads).
__global__ __launch_bounds__(256) void k_solve_scratch_init(ProblemDev hp, double shift, double* base,
                                                            const double* __restrict__ hup, int d, int dp)
{
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nthr = (size_t)gridDim.x * blockDim.x;
    if (tid == 0) {
        *reinterpret_cast<ProblemDev*>(base) = hp;
        base[kSolveSlotShift] = shift;
        *reinterpret_cast<int*>(base + kSolveSlotCtr) = 0;
    }
    if (tid < (size_t)kTraceStride)
        base[kSolveSlotTrace + tid] = 0.0;
    int* sync = reinterpret_cast<int*>(base + kSolveSlotSync);
    for (size_t e = tid; e < (size_t)mc_sync_words(dp); e += nthr)
        sync[e] = 0;
    double* W = base + kSolveOffW;
    const size_t total = (size_t)dp * dp;
    for (size_t e = tid; e < total; e += nthr) {
        const size_t i = e / dp, j = e - i * dp;  // column i, row j
        double v = 0.0;
        if (j >= i)
            v = ((int)j < d) ? hup[j * (j + 1) / 2 + i] : (i == j ? 1.0 : 0.0);
        W[e] = v;
    }
}

// The factorization flag of the status record to the caller's int.
__global__ void k_solve_status(const double* base, int* ok)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *ok = (base[kSolveSlotTrace + 3] == 1.0) ? 1 : 0;
}

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64];

// The calling thread's current device's copy of `scratch` (resolved once per
// device, with that device current).
static int solve_scratch(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_err("hipGetDevice", e);
        return FEDNL_E_DEVICE;
    }
    if (dev < 0 || dev >= 64) {
        set_err("fednl_newton_solve_noalloc: device index out of range");
        return FEDNL_E_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err("hipGetSymbolAddress (solve scratch)", e);
            return FEDNL_E_DEVICE;
        }
        e = raise_lds_limit((const void*)k_newton_mc, mc_lds_bytes(kSolveScratchMaxD));
        if (e != hipSuccess) {
            set_err("hipFuncSetAttribute (k_newton_mc)", e);
            return FEDNL_E_DEVICE;
        }
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
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
    o->line_search = 0;
    o->max_line_search_steps = 5;  // Args.h:300-306
    o->c_line_search = 0.5;
    o->gamma_line_search = 0.5;
    o->clients_per_round = 0;            // FedNL; > 0: FedNL-PP
    o->client_selection_seed = 1234u;    // bin_fednl_local_pp/include/Args.h:244
    o->hessian_form = FEDNL_HESS_AUTO;
}

const char* fednl_last_error(void) { return g_err.c_str(); }

int fednl_set_global_lr(void* handle, double global_lr)
{
    Engine* e = (Engine*)handle;
    if (!e) {
        set_err("fednl_set_global_lr: null handle");
        return FEDNL_E_MISSING_ARGUMENT;
    }
    if (e->o.algorithm != FEDNL_ALG_GD) {
        set_err("fednl_set_global_lr: the engine does not run GD");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    e->o.global_lr = global_lr;
    return sync_desc(e);
}

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
    int rc = enqueue_rounds(e, rounds);
    if (rc != FEDNL_OK)
        return rc;
    return fednl_sync(handle);
}

int fednl_batch_create(void* const* engines, int count, int use_graph, void** batch)
{
    if (!engines || count <= 0 || !batch) {
        set_err("fednl_batch_create: empty batch");
        return FEDNL_E_MISSING_ARGUMENT;
    }
    for (int i = 0; i < count; ++i)
        if (engines[i] && (((Engine*)engines[i])->pp_sel || ((Engine*)engines[i])->o.algorithm == FEDNL_ALG_GD)) {
            set_err("fednl_batch_create: FedNL-PP and GD engines run on their own");
            return FEDNL_E_UNSUPPORTED;
        }
    Batch* b = new Batch();
    b->use_graph = use_graph != 0;
    for (int i = 0; i < count; ++i) {
        Engine* e = (Engine*)engines[i];
        if (!e) {
            delete b;
            set_err("fednl_batch_create: null engine");
            return FEDNL_E_MISSING_ARGUMENT;
        }
        // the member's own queue (creation, earlier rounds) must be done
        if (hipStreamSynchronize(e->st) != hipSuccess || collect_timing(e) != FEDNL_OK) {
            delete b;
            set_err("fednl_batch_create: member stream");
            return FEDNL_E_DEVICE;
        }
        b->m.push_back(e);
    }
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        set_err("hipStreamCreate (batch)");
        return FEDNL_E_DEVICE;
    }
    int rc = batch_build(b);
    if (rc != FEDNL_OK) {
        free_batch(b);
        return rc;
    }
    *batch = b;
    return FEDNL_OK;
}

void fednl_batch_destroy(void* batch)
{
    Batch* b = (Batch*)batch;
    if (b && b->st)
        hipStreamSynchronize(b->st);
    free_batch(b);
}

int fednl_batch_enqueue(void* batch, int rounds)
{
    Batch* b = (Batch*)batch;
    if (!b)
        return FEDNL_E_MISSING_ARGUMENT;
    return batch_enqueue(b, rounds);
}

int fednl_batch_sync(void* batch)
{
    Batch* b = (Batch*)batch;
    if (!b)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(b->st));
    return FEDNL_OK;
}

int fednl_batch_run(void* batch, int rounds)
{
    int rc = fednl_batch_enqueue(batch, rounds);
    if (rc != FEDNL_OK)
        return rc;
    return fednl_batch_sync(batch);
}

size_t fednl_footprint(const fednl_opts* opts, int n, int m, int d)
{
    fednl_opts o;
    if (opts)
        o = *opts;
    else
        fednl_default_opts(&o);
    return engine_footprint(&o, n, m, d);
}

int fednl_serve(const fednl_request* reqs, int count, double hbm_budget_bytes, int rounds_per_step, int use_graph,
                fednl_result* out)
{
    if (!reqs || count <= 0 || !out || rounds_per_step <= 0) {
        set_err("fednl_serve: empty request list");
        return FEDNL_E_MISSING_ARGUMENT;
    }
    for (int i = 0; i < count; ++i)
        if (reqs[i].opts.clients_per_round > 0 || reqs[i].opts.algorithm == FEDNL_ALG_GD) {
            set_err("fednl_serve: FedNL-PP and GD requests are not batched");
            return FEDNL_E_UNSUPPORTED;
        }
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    struct Slot {
        int req;
        Engine* e;
        size_t bytes;
        int rounds;
    };
    std::vector<Slot> active;
    std::vector<int> waiting;  // arrived, not admitted (arrival order)
    int next = 0, done = 0;
    size_t used = 0;
    for (int i = 0; i < count; ++i) {
        memset(&out[i], 0, sizeof(fednl_result));
        out[i].status = 1;  // not finished
    }
    // requests are taken in arrival order
    std::vector<int> order(count);
    for (int i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    Batch* batch = nullptr;
    bool dirty = true;
    int rc = FEDNL_OK;
    while (done < count && rc == FEDNL_OK) {
        const double t = now_s();
        while (next < count && reqs[order[next]].arrival_s <= t)
            waiting.push_back(order[next++]);
        // admission by device-memory footprint, arrival order (no overtaking)
        while (!waiting.empty()) {
            const int q = waiting.front();
            const fednl_request& R = reqs[q];
            const size_t need = engine_footprint(&R.opts, R.n, R.m, R.d);
            if (!active.empty() && (double)(used + need) > hbm_budget_bytes)
                break;
            Engine* e = nullptr;
            const int crc = create_engine(&R.opts, R.n, R.m, R.d, R.features, R.labels, &e);
            out[q].admitted_s = now_s();
            if (crc != FEDNL_OK) {
                out[q].status = crc;
                out[q].finished_s = out[q].admitted_s;
                ++done;
                waiting.erase(waiting.begin());
                continue;
            }
            // a binary design's sparse lists are known once its data are read:
            // counted from here on (the budget can be passed by one request's lists)
            active.push_back({q, e, need + e->sp_bytes, 0});
            used += need + e->sp_bytes;
            waiting.erase(waiting.begin());
            dirty = true;
        }
        if (active.empty()) {
            if (next < count) {
                const double wait = reqs[order[next]].arrival_s - now_s();
                if (wait > 0)
                    std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
            }
            continue;
        }
        if (dirty) {
            if (batch) {
                hipStreamSynchronize(batch->st);
                free_batch(batch);
                batch = nullptr;
            }
            std::vector<void*> hs;
            for (auto& sl : active) hs.push_back(sl.e);
            if ((rc = fednl_batch_create(hs.data(), (int)hs.size(), use_graph, (void**)&batch)) != FEDNL_OK)
                break;
            dirty = false;
        }
        if ((rc = batch_enqueue(batch, rounds_per_step)) != FEDNL_OK)
            break;
        if (hipStreamSynchronize(batch->st) != hipSuccess) {
            set_err("hipStreamSynchronize (serve)");
            rc = FEDNL_E_DEVICE;
            break;
        }
        const double tnow = now_s();
        // retire members that reached the tolerance or their round budget
        for (size_t i = 0; i < active.size();) {
            Slot& sl = active[i];
            const fednl_request& R = reqs[sl.req];
            const int first = sl.rounds;
            sl.rounds += rounds_per_step;
            std::vector<double> tr((size_t)kTraceStride * rounds_per_step);
            hipMemcpy(tr.data(), sl.e->trace + (size_t)kTraceStride * first, sizeof(double) * tr.size(), hipMemcpyDeviceToHost);
            int hit = -1;
            for (int r = 0; r < rounds_per_step; ++r) {
                const double g = tr[(size_t)kTraceStride * r];
                if (g * g <= R.tol) {
                    hit = first + r;
                    break;
                }
            }
            const bool finished = (hit >= 0) || sl.rounds >= R.max_rounds;
            if (!finished) {
                ++i;
                continue;
            }
            fednl_result& o = out[sl.req];
            o.finished_s = tnow;
            o.rounds = sl.rounds;
            o.rounds_to_tol = hit;
            o.status = (hit >= 0) ? 0 : 2;
            o.grad_norm = tr[(size_t)kTraceStride * (rounds_per_step - 1)];
            if (R.x_out)
                hipMemcpy(R.x_out, sl.e->x, sizeof(double) * R.d, hipMemcpyDeviceToHost);
            used -= sl.bytes;
            free_engine(sl.e);
            active.erase(active.begin() + i);
            ++done;
            dirty = true;
        }
        if (dirty && batch) {
            free_batch(batch);
            batch = nullptr;
        }
    }
    if (batch) {
        hipStreamSynchronize(batch->st);
        free_batch(batch);
    }
    for (auto& sl : active) free_engine(sl.e);
    return rc;
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
    std::vector<double> tr((size_t)kTraceStride * R);
    HIP_OK(hipMemcpy(tr.data(), e->trace, sizeof(double) * tr.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < R; ++r) {
        if (grad_norm) grad_norm[r] = tr[(size_t)kTraceStride * r + 0];
        if (fval) fval[r] = tr[(size_t)kTraceStride * r + 1];
        if (x_norm) x_norm[r] = tr[(size_t)kTraceStride * r + 2];
        if (chol_ok) chol_ok[r] = tr[(size_t)kTraceStride * r + 3] == 1.0 ? 1 : 0;
        if (seconds)  // from the start of round 0 to the end of round r (device wall clock)
            seconds[r] = (tr[(size_t)kTraceStride * r + kTraceT1] - tr[kTraceT0]) / e->wclk_hz;
    }
    return FEDNL_OK;
}

int fednl_trace_clients(void* handle, int* participants, double* items, int* count)
{
    Engine* e = (Engine*)handle;
    if (!e || !count)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    const int R = e->rounds_enqueued;
    *count = R;
    if (R == 0)
        return FEDNL_OK;
    std::vector<double> tr((size_t)kTraceStride * R);
    HIP_OK(hipMemcpy(tr.data(), e->trace, sizeof(double) * tr.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < R; ++r) {
        if (participants) participants[r] = (int)tr[(size_t)kTraceStride * r + 5];
        if (items) items[r] = tr[(size_t)kTraceStride * r + 6];
    }
    return FEDNL_OK;
}

int fednl_pp_selection(void* handle, int* selected)
{
    Engine* e = (Engine*)handle;
    if (!e || !selected)
        return FEDNL_E_MISSING_ARGUMENT;
    if (!e->pp_sel) {
        set_err("fednl_pp_selection: not a FedNL-PP engine");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    HIP_OK(hipStreamSynchronize(e->st));
    HIP_OK(hipMemcpy(selected, e->pp_sel, sizeof(int) * e->n, hipMemcpyDeviceToHost));
    return FEDNL_OK;
}

int fednl_trace_line_search(void* handle, int* trials, int* count)
{
    Engine* e = (Engine*)handle;
    if (!e || !count)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    const int R = e->rounds_enqueued;
    *count = R;
    if (R == 0 || !trials)
        return FEDNL_OK;
    std::vector<double> tr((size_t)kTraceStride * R);
    HIP_OK(hipMemcpy(tr.data(), e->trace, sizeof(double) * tr.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < R; ++r) trials[r] = (int)tr[(size_t)kTraceStride * r + 4];
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
    // FedNL-PP: the full gradient at x_r (the direction estimate stays internal)
    if (grad) HIP_OK(hipMemcpy(grad, e->pp_tg ? e->pp_tg : e->g_avg, sizeof(double) * e->d, hipMemcpyDeviceToHost));
    if (lk) HIP_OK(hipMemcpy(lk, e->lk, sizeof(double) * e->n, hipMemcpyDeviceToHost));
    if (cnt) HIP_OK(hipMemcpy(cnt, e->cnt, sizeof(int) * e->n, hipMemcpyDeviceToHost));
    if (pos) HIP_OK(hipMemcpy(pos, e->pos, sizeof(uint32_t) * (size_t)e->n * e->k, hipMemcpyDeviceToHost));
    if (vals) HIP_OK(hipMemcpy(vals, e->vals, sizeof(double) * (size_t)e->n * e->k, hipMemcpyDeviceToHost));
    if (flat_idx) HIP_OK(hipMemcpy(flat_idx, e->idx, sizeof(uint32_t) * (size_t)e->n * e->k, hipMemcpyDeviceToHost));
    if (dx) HIP_OK(hipMemcpy(dx, e->dx, sizeof(double) * e->d, hipMemcpyDeviceToHost));
    return FEDNL_OK;
}

int fednl_last_delta(void* handle, double* delta)
{
    Engine* e = (Engine*)handle;
    if (!e || !delta)
        return FEDNL_E_MISSING_ARGUMENT;
    HIP_OK(hipStreamSynchronize(e->st));
    HIP_OK(hipMemcpy(delta, e->delta, sizeof(double) * (size_t)e->n * e->S, hipMemcpyDeviceToHost));
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
    return fednl_kernel_timing_mask(handle, enable ? ~0u : 0u);
}

int fednl_tuning(void* handle, int* margin_rows, int* hessian_chunk)
{
    Engine* e = (Engine*)handle;
    if (!e || !margin_rows)
        return FEDNL_E_MISSING_ARGUMENT;
    *margin_rows = e->margin_rows;
    if (hessian_chunk)
        *hessian_chunk = e->hess_chunk;
    return FEDNL_OK;
}

int fednl_hessian_form(void* handle)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    return e->hform;
}

int fednl_kernel_timing_mask(void* handle, unsigned mask)
{
    Engine* e = (Engine*)handle;
    if (!e)
        return FEDNL_E_MISSING_ARGUMENT;
    int rc = collect_timing(e);
    e->timing = mask != 0;
    e->timing_mask = mask;
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
        if (names) {
            // the Hessian and margin stages are named for the kernels that run them
            const char* hn = e->hform == FEDNL_HESS_SPARSE ? "k_hessian_sparse"
                             : e->hmode == 1               ? "k_hessian_f32"
                             : e->hmode == 2               ? "k_hessian_bf16"
                                                           : "k_hessian_f64";
            names[q] = q == KID_HESS ? hn
                       : (q == KID_MARGIN && e->hform == FEDNL_HESS_SPARSE) ? "k_margin_sparse"
                                                                           : kKernelNames[q];
        }
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
    if (select_keeps_k(compressor, k, S) && k > kMaxSelK) {
        set_err("fednl_select: k above the selection limit (32768)");
        return FEDNL_E_UNSUPPORTED;
    }
    workers = std::max(1, std::min(workers, nclients));
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double *ddelta = nullptr, *learned = nullptr, *lek_p = nullptr, *vals = nullptr;
    uint32_t *drng = nullptr, *dlek = nullptr, *sel = nullptr, *pos = nullptr, *idx = nullptr, *popped = nullptr;
    int *sel_cnt = nullptr, *hit = nullptr, *cnt = nullptr;
    unsigned char* scr = nullptr;  // wide selection scratch
    int rc = FEDNL_OK;
    auto cleanup = [&]() {
        void* ps[] = {ddelta, learned, lek_p, vals, drng, dlek, sel, pos, idx, popped, sel_cnt, hit, cnt, scr};
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
    const bool wide = select_wide(compressor, k, S);
    const bool big = nclients <= kBigRingClients;
    const size_t lds = select_lds(compressor, k, S, wide, big);
    const size_t lds_bm = emit_lds(S, wide);
    raise_lds_limit((const void*)k_select_heap<false, false, false>, lds);
    raise_lds_limit((const void*)k_select_heap<true, false, false>, lds);
    raise_lds_limit((const void*)k_select_randk<false, false>, lds);
    raise_lds_limit((const void*)k_select_heap<false, false, true>, lds);
    raise_lds_limit((const void*)k_select_heap<true, false, true>, lds);
    raise_lds_limit((const void*)k_emit<false>, lds_bm);
    const size_t scr_stride = wide ? select_scratch_stride(compressor, k, S) : 0;
    if (scr_stride && dalloc(&scr, (size_t)nclients * scr_stride) != hipSuccess) {
        set_err("hipMalloc (select scratch)");
        cleanup();
        return FEDNL_E_DEVICE;
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
    const int bm_words = (int)((S + 31) / 32);
    uint32_t* bm = nullptr;
    if (dalloc(&bm, (size_t)nclients * bm_words) != hipSuccess) {
        cleanup();
        return FEDNL_E_DEVICE;
    }
    // descriptor of the standalone call: delta as given, unit scales
    ProblemDev hp{};
    hp.n = nclients;
    hp.d = d;
    hp.k = k;
    hp.S = S;
    hp.lda = round_up(d, 4);
    hp.tiles_up = (compressor == FEDNL_TOPLEK) ? 1 : 0;
    hp.nparts = hp.tiles_up;
    hp.part = parts;
    hp.lek_delta = delta_param;
    hp.sel_ring = wide ? kMaxScanRing : scan_ring_slots(std::min(k, kMaxHeapK), S, big);
    hp.wide = wide ? 1 : 0;
    hp.sel_scr = scr;
    hp.sel_scr_stride = scr_stride;
    hp.comp = compressor;
    hp.workers = workers;
    hp.alpha = 1.0;
    hp.prescale = 1.0;
    hp.mult = 1.0;
    hp.delta = ddelta;
    hp.learned = learned;
    hp.sel = sel;
    hp.sel_cnt = sel_cnt;
    hp.lek_p = lek_p;
    hp.lek_hit = hit;
    hp.lek_popped = popped;
    hp.lek_state = dlek;
    hp.rng = drng;
    hp.seq_start = nullptr;
    hp.pos = pos;
    hp.idx = idx;
    hp.vals = vals;
    hp.cnt = cnt;
    hp.sel_bm = bm;
    hp.bm_words = bm_words;
    ProblemDev* dpd = nullptr;
    if (hipMalloc((void**)&dpd, sizeof(ProblemDev)) != hipSuccess ||
        hipMemcpyAsync(dpd, &hp, sizeof(ProblemDev), hipMemcpyHostToDevice, st) != hipSuccess) {
        set_err("descriptor upload (select)");
        if (dpd) hipFree(dpd);
        hipFree(bm);
        cleanup();
        return FEDNL_E_DEVICE;
    }
    const ProblemDev* pd = dpd;
    const int2* nomap = nullptr;
    const dim3 g(nclients), t(256);
    if (compressor == FEDNL_TOPK && !wide)
        hipLaunchKernelGGL((k_select_heap<false, false, false>), g, t, lds, st, pd, nomap);
    else if (compressor == FEDNL_TOPK)
        hipLaunchKernelGGL((k_select_heap<false, false, true>), g, t, lds, st, pd, nomap);
    else if (compressor == FEDNL_TOPLEK && !wide)
        hipLaunchKernelGGL((k_select_heap<true, false, false>), g, t, lds, st, pd, nomap);
    else if (compressor == FEDNL_TOPLEK)
        hipLaunchKernelGGL((k_select_heap<true, false, true>), g, t, lds, st, pd, nomap);
    else if (compressor == FEDNL_RANDK && !wide)
        hipLaunchKernelGGL((k_select_randk<false, false>), g, t, lds, st, pd, nomap);
    else if (compressor == FEDNL_RANDK)
        hipLaunchKernelGGL((k_select_randk<false, true>), g, t, lds, st, pd, nomap);
    else if (compressor == FEDNL_RANDSEQK || compressor == FEDNL_IDENTICAL)
        hipLaunchKernelGGL((k_select_window<false, false>), g, t, 0, st, pd, nomap);
    else {
        hipFree(dpd);
        hipFree(bm);
        cleanup();
        return FEDNL_E_UNSUPPORTED;
    }
    if (wide)
        hipLaunchKernelGGL(k_emit<true>, g, t, 0, st, pd, nomap);
    else
        hipLaunchKernelGGL(k_emit<false>, g, t, lds_bm, st, pd, nomap);
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
    hipFree(bm);
    hipFree(dpd);
    cleanup();
    return rc;
}

#ifdef FEDNL_PROBE
int fednl_probe_newton(void* handle, double* ny, double* nxt, double* W)
{
    Engine* e = (Engine*)handle;
    HIP_OK(hipStreamSynchronize(e->st));
    HIP_OK(hipMemcpy(ny, e->ny, sizeof(double) * e->dp, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(nxt, e->nxt, sizeof(double) * e->dp * kNB, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(W, e->W, sizeof(double) * e->dp * e->dp, hipMemcpyDeviceToHost));
    return 0;
}

int fednl_probe_read(unsigned long long* out, int count, int reset)
{
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_probe), sizeof(unsigned long long) * count) != hipSuccess)
        return -1;
    if (reset) {
        std::vector<unsigned long long> z(16384, 0ull);
        if (hipMemcpyToSymbol(HIP_SYMBOL(g_probe), z.data(), sizeof(unsigned long long) * z.size()) != hipSuccess)
            return -1;
    }
    return 0;
}
#endif

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

int fednl_probe_mfma_mixed(const float* a32, const float* b32, float* out32, const float* abf, const float* bbf,
                           float* outbf)
{
    float *d_a = nullptr, *d_o = nullptr;
    // inputs: a32[64], b32[64], abf[512], bbf[512]; outputs: 256 + 256 floats
    HIP_OK(hipMalloc((void**)&d_a, sizeof(float) * (64 + 64 + 512 + 512)));
    HIP_OK(hipMalloc((void**)&d_o, sizeof(float) * 512));
    hipMemcpy(d_a, a32, sizeof(float) * 64, hipMemcpyHostToDevice);
    hipMemcpy(d_a + 64, b32, sizeof(float) * 64, hipMemcpyHostToDevice);
    hipMemcpy(d_a + 128, abf, sizeof(float) * 512, hipMemcpyHostToDevice);
    hipMemcpy(d_a + 640, bbf, sizeof(float) * 512, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_mfma_probe_mixed, dim3(1), dim3(64), 0, 0, d_a, d_o);
    hipError_t le = hipGetLastError();
    if (le == hipSuccess)
        le = hipDeviceSynchronize();
    if (le == hipSuccess) {
        hipMemcpy(out32, d_o, sizeof(float) * 256, hipMemcpyDeviceToHost);
        hipMemcpy(outbf, d_o + 256, sizeof(float) * 256, hipMemcpyDeviceToHost);
    }
    hipFree(d_a);
    hipFree(d_o);
    if (le != hipSuccess) {
        set_err("mfma probe", le);
        return FEDNL_E_DEVICE;
    }
    return FEDNL_OK;
}

int fednl_hessians_at(void* handle, const double* x, double* out)
{
    Engine* e = (Engine*)handle;
    if (!e || !x || !out) {
        set_err("fednl_hessians_at: null argument");
        return FEDNL_E_WRONG_ARGUMENT;
    }
    double *dxv = nullptr, *dout = nullptr;
    HIP_OK(hipMalloc((void**)&dxv, sizeof(double) * e->d_pad));
    if (hipMalloc((void**)&dout, sizeof(double) * (size_t)e->n * e->S) != hipSuccess) {
        hipFree(dxv);
        set_err("hipMalloc (hessians_at)");
        return FEDNL_E_DEVICE;
    }
    std::vector<double> hx(e->d_pad, 0.0);
    std::copy(x, x + e->d, hx.begin());
    hipError_t err = hipStreamSynchronize(e->st);
    if (err == hipSuccess)
        err = hipMemcpyAsync(dxv, hx.data(), sizeof(double) * e->d_pad, hipMemcpyHostToDevice, e->st);
    ProblemDev* dpd = nullptr;
    ProblemDev hp = make_desc(e);
    hp.x = dxv;
    if (err == hipSuccess)
        err = hipMalloc((void**)&dpd, sizeof(ProblemDev));
    if (err == hipSuccess)
        err = hipMemcpy(dpd, &hp, sizeof(ProblemDev), hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        // the margin weights of x, then the Hessians in the engine's mode
        if (e->hform == FEDNL_HESS_SPARSE)
            launch_margin(e, false, dpd);
        else
            hipLaunchKernelGGL((k_margin<false, 4>), dim3(blocks_margin(hp)), dim3(256), sizeof(double) * e->d_pad,
                               e->st, (const ProblemDev*)dpd, (const int2*)nullptr);
        launch_hessian(e, true, dout, dpd);
        err = hipGetLastError();
    }
    if (err == hipSuccess)
        err = hipStreamSynchronize(e->st);
    if (err == hipSuccess)
        err = hipMemcpy(out, dout, sizeof(double) * (size_t)e->n * e->S, hipMemcpyDeviceToHost);
    hipFree(dxv);
    hipFree(dout);
    if (dpd)
        hipFree(dpd);
    if (err != hipSuccess) {
        set_err("fednl_hessians_at", err);
        return FEDNL_E_DEVICE;
    }
    return FEDNL_OK;
}

int fednl_newton_solve(int d, const double* H_up, double shift, const double* b, double* dx)
{
    if (d <= 0 || !H_up || !b || !dx)
        return FEDNL_E_MISSING_ARGUMENT;
    if (d > kMaxMcD) {
        set_err("fednl_newton_solve: d > 2048");
        return FEDNL_E_UNSUPPORTED;
    }
    const int dp = round_up(d, kNB);
    double *dsh = nullptr, *db = nullptr, *dW = nullptr, *dx_d = nullptr, *dtr = nullptr;
    double *dny = nullptr, *dninv = nullptr, *dnxt = nullptr;
    int *dctr = nullptr, *dsync = nullptr;
    ProblemDev* dpd = nullptr;
    auto cleanup = [&]() {
        void* ptrs[] = {dsh, db, dW, dx_d, dtr, dny, dninv, dnxt, dctr, dsync, dpd};
        for (void* q : ptrs)
            if (q)
                hipFree(q);
    };
    // W: the lower triangle of H (column i, rows j >= i), identity padding
    std::vector<double> w0((size_t)dp * dp, 0.0);
    for (int i = 0; i < dp; ++i)
        for (int j = i; j < dp; ++j)
            w0[(size_t)i * dp + j] = (j < d) ? H_up[(size_t)j * (j + 1) / 2 + i] : (i == j ? 1.0 : 0.0);
    hipError_t err = hipSuccess;
    auto step = [&](hipError_t r) {
        if (err == hipSuccess && r != hipSuccess)
            err = r;
    };
    step(dalloc(&dsh, 1));
    step(dalloc(&db, d));
    step(dalloc(&dW, (size_t)dp * dp));
    step(dalloc(&dx_d, d));
    step(dalloc(&dtr, kTraceStride));
    step(dalloc(&dny, dp));
    step(dalloc(&dninv, dp));
    step(dalloc(&dnxt, (size_t)dp * kNB));
    step(dalloc(&dctr, 1));
    step(dalloc(&dsync, mc_sync_words(dp)));
    step(dalloc(&dpd, 1));
    if (err == hipSuccess) {
        step(hipMemcpy(dsh, &shift, sizeof(double), hipMemcpyHostToDevice));
        step(hipMemcpy(db, b, sizeof(double) * d, hipMemcpyHostToDevice));
        step(hipMemcpy(dW, w0.data(), sizeof(double) * w0.size(), hipMemcpyHostToDevice));
        step(hipMemset(dctr, 0, sizeof(int)));
        step(hipMemset(dsync, 0, sizeof(int) * mc_sync_words(dp)));
        step(hipMemset(dtr, 0, sizeof(double) * kTraceStride));
        step(raise_lds_limit((const void*)k_newton_mc, mc_lds_bytes(dp)));
    }
    ProblemDev hp{};
    hp.d = d;
    hp.dp = dp;
    hp.trace_cap = 1;
    hp.update_x = 0;
    hp.shift = dsh;
    hp.nb = db;
    hp.W = dW;
    hp.x = nullptr;
    hp.dx = dx_d;
    hp.trace = dtr;
    hp.round_ctr = dctr;
    hp.nsync = dsync;
    hp.ny = dny;
    hp.ninv = dninv;
    hp.nxt = dnxt;
    if (err == hipSuccess)
        step(hipMemcpy(dpd, &hp, sizeof(ProblemDev), hipMemcpyHostToDevice));
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_newton_mc, dim3(mc_blocks(dp)), dim3(kMcThreads), mc_lds_bytes(dp), 0,
                           (const ProblemDev*)dpd, (const int2*)nullptr);
        step(hipGetLastError());
        step(hipDeviceSynchronize());
    }
    double tr[kTraceStride] = {0};
    if (err == hipSuccess)
        step(hipMemcpy(tr, dtr, sizeof(tr), hipMemcpyDeviceToHost));
    if (err == hipSuccess)
        step(hipMemcpy(dx, dx_d, sizeof(double) * d, hipMemcpyDeviceToHost));
    cleanup();
    if (err != hipSuccess) {
        set_err("fednl_newton_solve", err);
        return FEDNL_E_DEVICE;
    }
    if (tr[3] < 0.0) {
        set_err("fednl_newton_solve: a wait between workgroups expired");
        return FEDNL_E_DEVICE;
    }
    return tr[3] == 1.0 ? 1 : 0;
}

int fednl_newton_solve_noalloc(int d, const double* H_up, double shift, const double* b, double* dx, int* ok,
                               void* stream)
{
    if (d <= 0 || !H_up || !b || !dx || !ok) {
        set_err("fednl_newton_solve_noalloc: null argument or d <= 0");
        return FEDNL_E_MISSING_ARGUMENT;
    }
    if (d > kSolveScratchMaxD) {
        set_err("fednl_newton_solve_noalloc: d > 512 (the solve scratch holds a 512 x 512 workspace)");
        return FEDNL_E_UNSUPPORTED;
    }
    double* base = nullptr;
    int rc = solve_scratch(&base);
    if (rc != FEDNL_OK)
        return rc;
    const int dp = round_up(d, kNB);
    hipStream_t st = (hipStream_t)stream;
    ProblemDev hp{};
    hp.d = d;
    hp.dp = dp;
    hp.trace_cap = 1;
    hp.update_x = 0;
    hp.shift = base + kSolveSlotShift;
    hp.nb = b;
    hp.W = base + kSolveOffW;
    hp.x = nullptr;
    hp.dx = dx;
    hp.trace = base + kSolveSlotTrace;
    hp.round_ctr = reinterpret_cast<int*>(base + kSolveSlotCtr);
    hp.nsync = reinterpret_cast<int*>(base + kSolveSlotSync);
    hp.ny = base + kSolveOffNy;
    hp.ninv = base + kSolveOffNinv;
    hp.nxt = base + kSolveOffNxt;
    const int init_blocks = std::max(1, std::min(1024, (dp * dp + 255) / 256));
    hipLaunchKernelGGL(k_solve_scratch_init, dim3(init_blocks), dim3(256), 0, st, hp, shift, base, H_up, d, dp);
    hipLaunchKernelGGL(k_newton_mc, dim3(mc_blocks(dp)), dim3(kMcThreads), mc_lds_bytes(dp), st,
                       (const ProblemDev*)base, (const int2*)nullptr);
    hipLaunchKernelGGL(k_solve_status, dim3(1), dim3(64), 0, st, (const double*)base, ok);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_err("fednl_newton_solve_noalloc launch", e);
        return FEDNL_E_DEVICE;
    }
    return FEDNL_OK;
}

}  // extern "C"
