/*
 * fednl_oracle.c -- CPU restatement of dopt's single-node FedNL round
 * (burlachenkok/ufednl, dopt/bin_fednl_local), written for checking only.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load this library, and only as the checker. The
 * product path (ufednl_amd) never links, loads or calls it.
 *
 * Layout conventions (follow the reference):
 *   - d x d matrices are column-major with leading dimension
 *     LDA = roundUp(d, 4) (MatrixNMD.h:44-81, Vec4d batch of the AVX2 build);
 *     flat index of (i, j) is j*LDA + i (MatrixNMD.h:997-1000).
 *   - The upper triangle is enumerated column by column, i <= j
 *     (SpecialMathRoutinesForMatrix.h:35-64). Position p in that list is the
 *     "list position"; packed upper storage keeps element p at offset p.
 *
 * Floating point: build with -ffp-contract=off (no fused multiply-add), so
 * every a*b+c below rounds twice, as in the reference's no-FMA build.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Layout                                                                    */
/* ------------------------------------------------------------------------ */

/* MatrixNMD.h:44-81: rows rounded up to the Vec4d batch (4 doubles). */
ORC_API int orc_lda(int d) { return (d + 3) / 4 * 4; }

/* SpecialMathRoutinesForMatrix.h:35-64 */
ORC_API void orc_upper_list(int d, int lda, uint32_t *out)
{
    size_t w = 0;
    for (int j = 0; j < d; ++j)
        for (int i = 0; i <= j; ++i)
            out[w++] = (uint32_t)((size_t)j * lda + i);
}

/* ------------------------------------------------------------------------ */
/* Random generators                                                         */
/* ------------------------------------------------------------------------ */

/* RandomGenIntegerLinear.h:107-123 with A=0,B=2^32-1: raw LCG state. */
ORC_API uint32_t orc_lcg_int_next(uint32_t *s)
{
    *s = 214013u * (*s) + 2531011u;
    return *s;
}

/* RandomGenIntegerLinear.h:49-58 */
ORC_API double orc_lcg_int_unit(uint32_t *s)
{
    *s = 214013u * (*s) + 2531011u;
    return (double)(*s) * (1.0 / 4294967295.0);
}

/* RandomGenRealLinear.h:49-54, RandomGenRealLinear.cpp:5-8:
 * x = ((a*x + c) mod 2^32) mod (2^32 - 1). */
ORC_API double orc_lcg_real_unit(uint32_t *s)
{
    *s = (214013u * (*s) + 2531011u) % 4294967295u;
    return (double)(*s) * (1.0 / 4294967295.0);
}

/* SimpleMathRoutines.h:407-414 with TInteger = size_t */
static size_t round_nearest(double v)
{
    size_t t = (size_t)v;
    if (v - (double)t > 0.5)
        return t + 1;
    return t;
}

/* Shuffle.h:31-93: partial Fisher-Yates over the first k slots. */
ORC_API void orc_shuffle_partial(uint32_t *v, size_t n, size_t k, uint32_t *state)
{
    size_t rem = n - 1;
    for (size_t i = 0; i < k; ++i, --rem)
    {
        double u = orc_lcg_int_unit(state);
        size_t off = round_nearest(u * (double)rem);
        if (off != 0)
        {
            uint32_t t = v[i];
            v[i] = v[i + off];
            v[i + off] = t;
        }
    }
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------------ */
/* 4-ary min-heap exactly as MinHeap<4> (MinHeap.h:29-291)                    */
/*                                                                          */
/* The heap holds element ids; key[id] is the comparable value.             */
/* minHeapify keeps comparing against the value that last moved up          */
/* (MinHeap.h:259-288), which is reproduced as is.                          */
/* ------------------------------------------------------------------------ */

static double absv(double v) { return v > 0.0 ? v : -v; } /* SimpleMathRoutines.h:350-353 */

static void heapify(uint32_t *h, const double *key, size_t i, size_t n)
{
    double lowest_value = key[h[i]];
    for (;;)
    {
        size_t lowest = i;
        size_t first = 4 * i + 1;
        size_t bound = first + 4 < n ? first + 4 : n;
        for (size_t c = first; c < bound; ++c)
        {
            double cv = key[h[c]];
            if (lowest_value > cv)
            {
                lowest = c;
                lowest_value = cv;
            }
        }
        if (lowest != i)
        {
            uint32_t t = h[i];
            h[i] = h[lowest];
            h[lowest] = t;
            i = lowest;
        }
        else
            break;
    }
}

/* MinHeap.h:147-191 */
static void build_heap(uint32_t *h, const double *key, size_t n)
{
    size_t i = (n + 4 - 2) / 4;
    for (;;)
    {
        heapify(h, key, i, n);
        if (i == 0)
            break;
        --i;
    }
}

/* MinHeap.h:92-103 */
static void heap_insert(uint32_t *h, const double *key, size_t *n, uint32_t id)
{
    size_t i = (*n)++;
    h[i] = id;
    while (i != 0 && key[h[(i - 1) / 4]] > key[h[i]])
    {
        uint32_t t = h[(i - 1) / 4];
        h[(i - 1) / 4] = h[i];
        h[i] = t;
        i = (i - 1) / 4;
    }
}

/* MinHeap.h:125-141 */
static uint32_t heap_extract(uint32_t *h, const double *key, size_t *n)
{
    uint32_t res = h[0];
    size_t sz = *n;
    uint32_t t = h[0];
    h[0] = h[sz - 1];
    h[sz - 1] = t;
    heapify(h, key, 0, sz - 1);
    *n = sz - 1;
    return res;
}

/* Heap scan of getTopKFromUpperDiagonalPart<true> (SpecialMathRoutinesForMatrix.h:276-327)
 * on a key array indexed by list position. Leaves k list positions in h
 * (heap order, not sorted). */
static void topk_scan(const double *key, size_t S, size_t k, uint32_t *h)
{
    for (size_t p = 0; p < k; ++p)
        h[p] = (uint32_t)p;
    build_heap(h, key, k);
    double minv = key[h[0]];
    for (size_t p = k; p < S; ++p)
    {
        if (key[p] > minv)
        {
            h[0] = (uint32_t)p; /* increaseMinimumInMinHeap, MinHeap.h:78-85 */
            heapify(h, key, 0, k);
            minv = key[h[0]];
        }
    }
}

/* TopK on packed upper values (list order). Output: sorted list positions.
 * Returns the count (k, or S when k >= S). */
ORC_API int orc_topk_packed(const double *up, size_t S, size_t k, uint32_t *out_pos)
{
    if (k >= S)
    {
        for (size_t p = 0; p < S; ++p)
            out_pos[p] = (uint32_t)p;
        return (int)S;
    }
    double *key = (double *)malloc(sizeof(double) * S);
    for (size_t p = 0; p < S; ++p)
        key[p] = absv(up[p]);
    topk_scan(key, S, k, out_pos);
    qsort(out_pos, k, sizeof(uint32_t), cmp_u32);
    free(key);
    return (int)k;
}

/* Packed upper from a full d x LDA matrix and back. */
ORC_API void orc_pack_upper(const double *mat, int d, int lda, double *up)
{
    size_t w = 0;
    for (int j = 0; j < d; ++j)
        for (int i = 0; i <= j; ++i)
            up[w++] = mat[(size_t)j * lda + i];
}

/* Flat index (j*LDA+i) of list position p. */
static uint32_t flat_of_pos(size_t p, int lda)
{
    size_t j = (size_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) / 2.0);
    while (j * (j + 1) / 2 > p)
        --j;
    while ((j + 1) * (j + 2) / 2 <= p)
        ++j;
    size_t i = p - j * (j + 1) / 2;
    return (uint32_t)(j * (size_t)lda + i);
}

ORC_API uint32_t orc_flat_of_pos(size_t p, int lda) { return flat_of_pos(p, lda); }

/* TopK on a full d x LDA matrix; output sorted flat indices (as the reference). */
ORC_API int orc_topk(const double *mat, int d, int lda, int k, uint32_t *out)
{
    size_t S = (size_t)d * (d + 1) / 2;
    double *up = (double *)malloc(sizeof(double) * S);
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * S);
    orc_pack_upper(mat, d, lda, up);
    int cnt = orc_topk_packed(up, S, (size_t)k, pos);
    for (int q = 0; q < cnt; ++q)
        out[q] = flat_of_pos(pos[q], lda);
    free(up);
    free(pos);
    return cnt;
}

/* frobeniusNormSquareForSymmetricMatrixFromUpPart for VectorNDRaw_d with the
 * AVX2 code path (MatrixNMD_SIMD_double.h:1067-1156): two Vec4d accumulators
 * over 8-element chunks of each column's strict upper part, a scalar
 * remainder, a scalar diagonal sum; horizontal_add(Vec4d) is
 * (a0+a2)+(a1+a3) (vectorclass vectorf256.h:1967, vectorf128.h:2016-2028).
 * Input: packed upper values. */
ORC_API double orc_frob_upper_sqr_packed(const double *up, int d)
{
    double c[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    double rest = 0.0, diag = 0.0;
    size_t col = 0;
    for (int j = 0; j < d; ++j)
    {
        const double *a = up + col;
        size_t items = (size_t)j / 8 * 8;
        size_t i = 0;
        for (; i < items; i += 8)
            for (int u = 0; u < 2; ++u)
                for (int l = 0; l < 4; ++l)
                {
                    double v = a[i + 4 * u + l];
                    double sq = v * v;
                    c[u][l] = c[u][l] + sq;
                }
        for (; i < (size_t)j; ++i)
        {
            double v = a[i];
            double sq = v * v;
            rest = rest + sq;
        }
        double dv = a[j];
        double dsq = dv * dv;
        diag = diag + dsq;
        col += (size_t)j + 1;
    }
    double out = 0.0;
    for (int u = 0; u < 2; ++u)
    {
        double lo = c[u][0] + c[u][2];
        double hi = c[u][1] + c[u][3];
        double h = lo + hi;
        out = out + h;
    }
    out = out + rest;
    return (out + out) + diag;
}

ORC_API double orc_frob_upper_packed(const double *up, int d)
{
    return sqrt(orc_frob_upper_sqr_packed(up, d));
}

/* getTopLEKFromUpperDiagonalPart<true> (SpecialMathRoutinesForMatrix.h:377-511)
 * on packed upper values. rng: state of the worker's RandomGenRealLinear
 * stream (thread_local, seed 1234, :383). delta = k / S (the TopLEK worker,
 * LocalWorkerForFedNL_TopLEK.h:19,58). Output sorted list positions; returns
 * the count (0..k). */
ORC_API int orc_toplek_packed(const double *up, int d, size_t k, double delta, uint32_t *rng,
                              uint32_t *out_pos)
{
    size_t S = (size_t)d * (d + 1) / 2;
    double one_minus_delta = 1.0 - delta;
    double fro2 = orc_frob_upper_sqr_packed(up, d);
    double target = one_minus_delta * fro2;
    double eps = 1e-1 * fro2;
    if (k >= S)
    {
        for (size_t p = 0; p < S; ++p)
            out_pos[p] = (uint32_t)p;
        return (int)S;
    }
    double *key = (double *)malloc(sizeof(double) * S);
    uint32_t *h = (uint32_t *)malloc(sizeof(uint32_t) * (k + 1));
    for (size_t p = 0; p < S; ++p)
        key[p] = absv(up[p]);
    topk_scan(key, S, k, h);
    size_t n = k;

    double numerator = fro2;
    for (size_t q = 0; q < n; ++q)
    {
        size_t p = h[q];
        double item = up[p];
        /* the diagonal test of getTranspoedIndexFromFlatternIndex: list
         * position p is diagonal iff it is the last one of its column */
        uint32_t f = flat_of_pos(p, 4 * ((d + 3) / 4));
        size_t lda = (size_t)(4 * ((d + 3) / 4));
        size_t jc = f / lda, ir = f - jc * lda;
        if (jc == ir)
            numerator -= item * item;
        else
            numerator -= 2.0 * item * item;
    }
    do
    {
        double prev = numerator;
        uint32_t popped = heap_extract(h, key, &n);
        double item = up[popped];
        uint32_t f = flat_of_pos(popped, 4 * ((d + 3) / 4));
        size_t lda = (size_t)(4 * ((d + 3) / 4));
        size_t jc = f / lda, ir = f - jc * lda;
        if (jc == ir)
            numerator += item * item;
        else
            numerator += 2.0 * item * item;
        double next = numerator;
        if (target + eps > prev && target < next + eps)
        {
            double p = (target - prev + eps) / (next - prev + eps);
            if (orc_lcg_real_unit(rng) > p)
                heap_insert(h, key, &n, popped);
            break;
        }
    } while (n != 0);

    memcpy(out_pos, h, n * sizeof(uint32_t));
    qsort(out_pos, n, sizeof(uint32_t), cmp_u32);
    free(key);
    free(h);
    return (int)n;
}

/* generateRandKItemsInUpperTriangularPart (SpecialMathRoutinesForMatrix.h:95-107)
 * in list positions (the list is monotone in flat index). */
ORC_API void orc_randk_pos(uint32_t *state, size_t S, size_t k, uint32_t *out_pos)
{
    uint32_t *v = (uint32_t *)malloc(sizeof(uint32_t) * S);
    for (size_t p = 0; p < S; ++p)
        v[p] = (uint32_t)p;
    orc_shuffle_partial(v, S, k, state);
    memcpy(out_pos, v, k * sizeof(uint32_t));
    qsort(out_pos, k, sizeof(uint32_t), cmp_u32);
    free(v);
}

/* generateRandSeqKItemsInUpperTriangularPart (SpecialMathRoutinesForMatrix.h:137-169):
 * cyclic window of k list positions starting at generateInteger() % S (not sorted). */
ORC_API void orc_randseqk_pos(uint32_t *state, size_t S, size_t k, uint32_t *out_pos)
{
    uint32_t r = orc_lcg_int_next(state);
    size_t start = (size_t)r % S;
    for (size_t q = 0; q < k; ++q)
    {
        size_t p = start + q;
        if (p >= S)
            p -= S;
        out_pos[q] = (uint32_t)p;
    }
}

/* ------------------------------------------------------------------------ */
/* L2-regularised logistic regression oracle (logistic_regression.h)         */
/* feat: m x d row-major sample features, labels +-1.                       */
/* Atr[i,s] = b_s * a_s[i] (ctor, :26-76).                                  */
/* ------------------------------------------------------------------------ */

static double atr(const double *feat, const double *labels, int d, int s, int i)
{
    double a = feat[(size_t)s * d + i];
    return labels[s] == 1.0 ? a : a * -1.0;
}

/* Margin, sigmoid, gradient, f and packed-upper Hessian of one client.
 * Sample order follows the reference; sums run sequentially.
 * hess_up (S values, list order) includes +lambda on the diagonal
 * (evaluateHessian<true,4>, :365-547 then addToAllDiagonalEntries). Any
 * output pointer may be NULL. */
ORC_API void orc_client_eval(const double *feat, const double *labels, int d, int m, double lambda,
                             const double *x, double *margin, double *sigm, double *grad,
                             double *hess_up, double *fval, double *dweights)
{
    double inv_m = 1.0 / (double)m;
    double minus_inv_m = -inv_m;
    double *t = (double *)malloc(sizeof(double) * m);
    double *sg = (double *)malloc(sizeof(double) * m);
    double *D = (double *)malloc(sizeof(double) * m);
    for (int s = 0; s < m; ++s)
    {
        double acc = 0.0;
        for (int i = 0; i < d; ++i)
        {
            double pr = atr(feat, labels, d, s, i) * x[i];
            acc = acc + pr;
        }
        t[s] = acc;
        /* VectorND_Raw_SIMD_double.h:1596-1650 */
        sg[s] = 1.0 / (1.0 + exp(-acc));
        /* :398-404 */
        double z1 = 1.0 - sg[s];
        double z2 = z1 * sg[s];
        D[s] = inv_m * z2;
    }
    if (margin)
        memcpy(margin, t, sizeof(double) * m);
    if (sigm)
        memcpy(sigm, sg, sizeof(double) * m);
    if (dweights)
        memcpy(dweights, D, sizeof(double) * m);
    if (grad)
    {
        /* :148-161 -> scaledDifferenceWithEye: g_s = (1 - sigma_s) * (-1/m);
         * gradient = lambda*x + Atr*g */
        for (int i = 0; i < d; ++i)
            grad[i] = lambda * x[i];
        for (int s = 0; s < m; ++s)
        {
            double gs = (1.0 - sg[s]) * minus_inv_m;
            for (int i = 0; i < d; ++i)
            {
                double pr = atr(feat, labels, d, s, i) * gs;
                grad[i] = grad[i] + pr;
            }
        }
    }
    if (fval)
    {
        /* :119-123 -> logisticUnweightedLossFromMarginSigmoid: mean(-log sigma) */
        double acc = 0.0;
        for (int s = 0; s < m; ++s)
            acc = acc - log(sg[s]);
        double xx = 0.0;
        for (int i = 0; i < d; ++i)
            xx = xx + x[i] * x[i];
        *fval = acc / (double)m + (lambda / 2.0) * xx;
    }
    if (hess_up)
    {
        /* per element H[i,j] = sum_s Atr[i,s] * (D_s * Atr[j,s]), s ascending */
        size_t S = (size_t)d * (d + 1) / 2;
        memset(hess_up, 0, sizeof(double) * S);
        double *col = (double *)malloc(sizeof(double) * d);
        for (int s = 0; s < m; ++s)
        {
            for (int i = 0; i < d; ++i)
                col[i] = atr(feat, labels, d, s, i);
            size_t w = 0;
            for (int j = 0; j < d; ++j)
            {
                double vj = D[s] * col[j];
                for (int i = 0; i <= j; ++i)
                {
                    double pr = col[i] * vj;
                    hess_up[w] = hess_up[w] + pr;
                    ++w;
                }
            }
        }
        for (int j = 0; j < d; ++j)
        {
            size_t dpos = (size_t)j * (j + 1) / 2 + j;
            hess_up[dpos] = hess_up[dpos] + lambda;
        }
        free(col);
    }
    free(t);
    free(sg);
    free(D);
}

/* ------------------------------------------------------------------------ */
/* Newton system: Cholesky (CholeskyFactorization.h:153-272) of             */
/* H_upper + shift*I, then forward/backward substitution                    */
/* (ElementarySolvers.h:187-226, 72-116). Upper factor U (= Ltr's upper      */
/* part) in packed upper storage. Returns 1 on success, 0 on failure.        */
/* ------------------------------------------------------------------------ */

#define UP(i, j) ((size_t)(j) * ((j) + 1) / 2 + (i)) /* i <= j */

ORC_API int orc_cholesky_packed(const double *H_up, int d, double shift, double *U)
{
    for (int i = 0; i < d; ++i)
    {
        for (int j = 0; j < i; ++j)
        {
            double sq = 0.0;
            for (int q = 0; q < j; ++q)
                sq = sq + U[UP(q, i)] * U[UP(q, j)];
            double ljj = U[UP(j, j)];
            U[UP(j, i)] = (1.0 / ljj) * (H_up[UP(j, i)] - sq);
        }
        double sq = 0.0;
        for (int q = 0; q < i; ++q)
            sq = sq + U[UP(q, i)] * U[UP(q, i)];
        double r = (H_up[UP(i, i)] + shift) - sq;
        if (r < 0.0)
            return 0;
        U[UP(i, i)] = sqrt(r);
    }
    return 1;
}

ORC_API void orc_tri_solve_packed(const double *U, int d, const double *b, double *x)
{
    double *y = (double *)malloc(sizeof(double) * d);
    for (int i = 0; i < d; ++i)
    {
        double acc = 0.0;
        for (int q = 0; q < i; ++q)
            acc = acc + U[UP(q, i)] * y[q];
        y[i] = (b[i] - acc) / U[UP(i, i)];
    }
    for (int i = d - 1; i >= 0; --i)
    {
        double acc = 0.0;
        for (int q = i + 1; q < d; ++q)
            acc = acc + U[UP(i, q)] * x[q];
        x[i] = (y[i] - acc) / U[UP(i, i)];
    }
    free(y);
}

/* ------------------------------------------------------------------------ */
/* FedNL round engine (TrainLoop.h:1054-1698, forcesequpdate semantics)      */
/* ------------------------------------------------------------------------ */

enum { ORC_IDENTICAL = 0, ORC_RANDK = 1, ORC_TOPK = 2, ORC_RANDSEQK = 3, ORC_TOPLEK = 4, ORC_NATURAL = 5 };

/* Natural compressor of dopt (MatrixNMD.h:2009-2029, applyNaturalCompressorNaive):
 * the mantissa bits are cleared, sign and exponent kept (the value the
 * 24-bit pair packing of LocalWorkerForFedNL_Natural.h:80-126 carries). */
ORC_API double orc_natural(double v)
{
    uint64_t b;
    memcpy(&b, &v, 8);
    b &= 0xFFF0000000000000ull;
    memcpy(&v, &b, 8);
    return v;
}

/* Gauss-Jordan elimination with partial pivoting to reduced row-echelon
 * form (GaussEliminationSolvers.h:27-120, zero tolerance): for each column,
 * the row with the first largest |a(i, col)| at or below the current row is
 * swapped up, the row and b are divided by the pivot (b by division, the row
 * by the reciprocal), and the column is eliminated from every other row.
 * M: d x d row-major (not modified), x[pivot column of row i] = b[i]. */
ORC_API void orc_gauss_solve(const double *M, const double *b, int d, double *x)
{
    double *a = (double *)malloc(sizeof(double) * (size_t)d * d);
    double *bb = (double *)malloc(sizeof(double) * d);
    int *piv = (int *)malloc(sizeof(int) * d);
    memcpy(a, M, sizeof(double) * (size_t)d * d);
    memcpy(bb, b, sizeof(double) * d);
    int row = 0, col = 0, np = 0;
    while (row < d && col < d)
    {
        double best = fabs(a[(size_t)row * d + col]);
        int br = row;
        for (int i = row + 1; i < d; ++i)
        {
            const double v = fabs(a[(size_t)i * d + col]);
            if (v > best)
            {
                br = i;
                best = v;
            }
        }
        if (br != row)
        {
            for (int j = 0; j < d; ++j)
            {
                const double t = a[(size_t)row * d + j];
                a[(size_t)row * d + j] = a[(size_t)br * d + j];
                a[(size_t)br * d + j] = t;
            }
            const double t = bb[row];
            bb[row] = bb[br];
            bb[br] = t;
        }
        if (best < 0.0)
        {
            col += 1;
            continue;
        }
        bb[row] = bb[row] / a[(size_t)row * d + col];
        const double inv = 1.0 / a[(size_t)row * d + col];
        for (int j = 0; j < d; ++j)
            a[(size_t)row * d + j] = a[(size_t)row * d + j] * inv;
        for (int i = 0; i < d; ++i)
        {
            if (i == row)
                continue;
            const double kq = a[(size_t)i * d + col];
            bb[i] = bb[i] - bb[row] * kq;
            for (int j = 0; j < d; ++j)
                a[(size_t)i * d + j] = a[(size_t)i * d + j] + a[(size_t)row * d + j] * (-kq);
            a[(size_t)i * d + col] = 0.0;
        }
        piv[np++] = col;
        ++row;
        ++col;
    }
    for (int i = 0; i < d; ++i)
        x[i] = 0.0;
    for (int i = 0; i < np; ++i)
        x[piv[i]] = bb[i];
    free(a); free(bb); free(piv);
}

typedef struct
{
    int n, m, d, lda, k, compressor, option_b, workers;
    size_t S;
    double lambda, alpha, prescale, inv_n, mult, delta_toplek;
    const double *feat, *labels; /* borrowed */
    double *x;                   /* d */
    double *learned;             /* n*S */
    double *H;                   /* S */
    double *H0;                  /* S: H after creation; the master's full H keeps it as its
                                    lower triangle (only upper entries are ever updated) */
    uint32_t *rng;               /* n client RandK/SeqK generators, seed 123+c */
    uint32_t *lek;               /* workers TopLEK streams, seed 1234 */
    /* scratch */
    double *grads, *lks, *fis, *hx, *delta, *vals;
    uint32_t *pos;
    int *cnt;
    /* FedNL-LS (TrainLoop.h:344-355, 1546-1640): Armijo backtracking on the
     * Newton direction from t = 1 */
    int ls, ls_max, ls_count;
    double ls_c, ls_gamma;
    /* FedNL-PP (bin_fednl_local_pp/include/TrainLoop.h): tau clients per
     * round, the client sampler, each client's last direction and Lk, the
     * master's running direction average and Lk average */
    int pp_tau;
    uint32_t pp_rng;
    double *pp_dir, *pp_lk, *pp_g, *pp_gmsg, *pp_lkmsg;
    double pp_li;
    int *pp_sel;
} orc_fednl;

ORC_API void orc_fednl_set_line_search(void *h, int on, double c, double gamma, int max_steps)
{
    orc_fednl *f = (orc_fednl *)h;
    f->ls = on;
    f->ls_c = c;
    f->ls_gamma = gamma;
    f->ls_max = max_steps;
}

ORC_API int orc_fednl_line_search_count(void *h) { return ((orc_fednl *)h)->ls_count; }

ORC_API void *orc_fednl_create(const double *feat, const double *labels, int n, int m, int d,
                               double lambda, int compressor, int k, double alpha, int option_b,
                               int workers)
{
    orc_fednl *f = (orc_fednl *)calloc(1, sizeof(orc_fednl));
    f->n = n; f->m = m; f->d = d; f->lda = orc_lda(d);
    f->k = k; f->compressor = compressor; f->option_b = option_b;
    f->workers = workers < 1 ? 1 : (workers > n ? n : workers);
    f->S = (size_t)d * (d + 1) / 2;
    f->lambda = lambda; f->alpha = alpha;
    f->inv_n = 1.0 / (double)n;                /* TrainLoop.h:562 */
    f->prescale = f->inv_n * alpha;            /* TrainLoop.h:564 */
    f->mult = (double)f->S / (double)k;        /* LocalWorkerForFedNL_RandK.h:90 */
    f->delta_toplek = (double)k * (1.0 / (double)f->S);
    f->feat = feat; f->labels = labels;
    size_t S = f->S;
    f->x = (double *)calloc(d, sizeof(double));
    f->learned = (double *)calloc((size_t)n * S, sizeof(double));
    f->H = (double *)calloc(S, sizeof(double));
    f->rng = (uint32_t *)malloc(sizeof(uint32_t) * n);
    f->lek = (uint32_t *)malloc(sizeof(uint32_t) * f->workers);
    f->grads = (double *)malloc(sizeof(double) * (size_t)n * d);
    f->lks = (double *)malloc(sizeof(double) * n);
    f->fis = (double *)malloc(sizeof(double) * n);
    f->hx = (double *)malloc(sizeof(double) * S);
    f->delta = (double *)malloc(sizeof(double) * S);
    f->vals = (double *)malloc(sizeof(double) * (size_t)n * (k < (int)S ? S : (size_t)k));
    f->pos = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n * (S > (size_t)k ? S : (size_t)k));
    f->cnt = (int *)malloc(sizeof(int) * n);
    for (int c = 0; c < n; ++c)
        f->rng[c] = (uint32_t)(123 + c);  /* TrainLoop.h:632 */
    for (int w = 0; w < f->workers; ++w)
        f->lek[w] = 1234u;                /* RandomGenRealLinear default seed */
    /* learned_c = H_c(x0); H = (sum_c learned_c) / n  (TrainLoop.h:639-643,732-735) */
    for (int c = 0; c < n; ++c)
    {
        double *L = f->learned + (size_t)c * S;
        orc_client_eval(feat + (size_t)c * m * d, labels + (size_t)c * m, d, m, lambda, f->x,
                        NULL, NULL, NULL, L, NULL, NULL);
        for (size_t p = 0; p < S; ++p)
            f->H[p] = f->H[p] + L[p];
    }
    for (size_t p = 0; p < S; ++p)
        f->H[p] = f->H[p] * f->inv_n;
    f->H0 = (double *)malloc(sizeof(double) * S);
    memcpy(f->H0, f->H, sizeof(double) * S);
    return f;
}

ORC_API void orc_fednl_destroy(void *h)
{
    orc_fednl *f = (orc_fednl *)h;
    if (!f)
        return;
    free(f->x); free(f->learned); free(f->H); free(f->H0); free(f->rng); free(f->lek);
    free(f->grads); free(f->lks); free(f->fis); free(f->hx); free(f->delta);
    free(f->vals); free(f->pos); free(f->cnt);
    free(f->pp_dir); free(f->pp_lk); free(f->pp_g); free(f->pp_gmsg); free(f->pp_lkmsg); free(f->pp_sel);
    free(f);
}

/* State access (x: d, learned: n*S packed, H: S packed, rng: n, lek: workers). */
ORC_API void orc_fednl_get_state(void *h, double *x, double *learned, double *H, uint32_t *rng,
                                 uint32_t *lek)
{
    orc_fednl *f = (orc_fednl *)h;
    if (x) memcpy(x, f->x, sizeof(double) * f->d);
    if (learned) memcpy(learned, f->learned, sizeof(double) * f->n * f->S);
    if (H) memcpy(H, f->H, sizeof(double) * f->S);
    if (rng) memcpy(rng, f->rng, sizeof(uint32_t) * f->n);
    if (lek) memcpy(lek, f->lek, sizeof(uint32_t) * f->workers);
}

ORC_API void orc_fednl_set_state(void *h, const double *x, const double *learned, const double *H,
                                 const uint32_t *rng, const uint32_t *lek)
{
    orc_fednl *f = (orc_fednl *)h;
    if (x) memcpy(f->x, x, sizeof(double) * f->d);
    if (learned) memcpy(f->learned, learned, sizeof(double) * f->n * f->S);
    if (H) memcpy(f->H, H, sizeof(double) * f->S);
    if (rng) memcpy(f->rng, rng, sizeof(uint32_t) * f->n);
    if (lek) memcpy(f->lek, lek, sizeof(uint32_t) * f->workers);
}

/* The client step of a round (LocalWorkerForFedNL_*.h): H_c(x), Delta,
 * Lk, selection, message values and the learned update; f->hx keeps H_c(x). */
static void client_step(orc_fednl *f, int c, int want_f)
{
    const int m = f->m, d = f->d;
    const size_t S = f->S, k = (size_t)f->k;
    const size_t cap = S > k ? S : k;
    {
        const double *feat = f->feat + (size_t)c * m * d;
        const double *lab = f->labels + (size_t)c * m;
        double *L = f->learned + (size_t)c * S;
        double fi = 0.0;
        orc_client_eval(feat, lab, d, m, f->lambda, f->x, NULL, NULL, f->grads + (size_t)c * d, f->hx,
                        want_f ? &fi : NULL, NULL);
        f->fis[c] = fi;
        for (size_t p = 0; p < S; ++p)
            f->delta[p] = f->hx[p] - L[p];
        f->lks[c] = orc_frob_upper_packed(f->delta, d);
        uint32_t *pc = f->pos + (size_t)c * cap;
        double *vc = f->vals + (size_t)c * cap;
        double scale = 1.0;
        int ck;
        switch (f->compressor)
        {
        case ORC_TOPK:
            ck = orc_topk_packed(f->delta, S, k, pc);
            break;
        case ORC_TOPLEK:
            ck = orc_toplek_packed(f->delta, d, k, f->delta_toplek, &f->lek[c % f->workers], pc);
            break;
        case ORC_RANDK:
            orc_randk_pos(&f->rng[c], S, k, pc);
            ck = (int)k;
            scale = f->mult;
            break;
        case ORC_RANDSEQK:
            orc_randseqk_pos(&f->rng[c], S, k, pc);
            ck = (int)k;
            scale = f->mult;
            break;
        default:
            for (size_t p = 0; p < S; ++p)
                pc[p] = (uint32_t)p;
            ck = (int)S;
            break;
        }
        f->cnt[c] = ck;
        const int nat = (f->compressor == ORC_NATURAL);
        for (int q = 0; q < ck; ++q)
        {
            double item = f->delta[pc[q]];
            if (nat)
                item = orc_natural(item);
            if (scale != 1.0)
                item = item * scale;
            vc[q] = item * f->prescale;
        }
        for (int q = 0; q < ck; ++q)
        {
            double item = f->delta[pc[q]];
            if (nat)
                item = orc_natural(item);
            if (scale != 1.0)
                item = item * scale;
            double upd = f->alpha * item;
            L[pc[q]] = L[pc[q]] + upd;
        }
    }
}

/* One round. Outputs (any may be NULL):
 *   grad[d]   averaged gradient at x_r,
 *   lk[n]     per-client Lk,
 *   cnt[n]    selected counts, pos[n*k] selected list positions (sorted,
 *             RandSeqK in window order), vals[n*k] values sent to the master,
 *   dx[d]     Newton step, stats[3] = (||grad||, f(x_r) or 0, ||x_r||).
 * Returns 1 on success, 0 if the Cholesky factorization failed. */
ORC_API int orc_fednl_round(void *h, int tracking, double *grad, double *lk, int *cnt, uint32_t *pos,
                            double *vals, double *dx, double *stats)
{
    orc_fednl *f = (orc_fednl *)h;
    int n = f->n, m = f->m, d = f->d;
    size_t S = f->S, k = (size_t)f->k;
    size_t cap = S > k ? S : k;
    double xx = 0.0;
    for (int i = 0; i < d; ++i)
        xx = xx + f->x[i] * f->x[i];
    double xnorm = sqrt(xx);

    for (int c = 0; c < n; ++c)
        client_step(f, c, tracking || f->ls);

    /* master: clients in order 0..n-1 */
    double *g = (double *)calloc(d, sizeof(double));
    double lsum = 0.0, fagg = 0.0;
    for (int c = 0; c < n; ++c)
    {
        const double *gc = f->grads + (size_t)c * d;
        for (int i = 0; i < d; ++i)
            g[i] = g[i] + gc[i];
        const uint32_t *pc = f->pos + (size_t)c * cap;
        const double *vc = f->vals + (size_t)c * cap;
        for (int q = 0; q < f->cnt[c]; ++q)
            f->H[pc[q]] = f->H[pc[q]] + vc[q];
        lsum = lsum + f->lks[c];
        fagg = fagg + f->fis[c];
    }
    for (int i = 0; i < d; ++i)
        g[i] = g[i] * f->inv_n;
    lsum = lsum * f->inv_n;
    fagg = fagg * f->inv_n;

    double shift = f->option_b ? lsum : f->lambda; /* option a: mu = lambda */
    double *U = (double *)malloc(sizeof(double) * S);
    double *step = (double *)calloc(d, sizeof(double));
    int ok = orc_cholesky_packed(f->H, d, shift, U);
    if (ok)
        orc_tri_solve_packed(U, d, g, step);
    else
    {
        /* TrainLoop.h:1410-1420 / 1476-1486: "SST FACTORIZATION FAILED", then
         * dx = gausEleminationSolver(H + shift I, g) on the master's full H:
         * the current upper triangle and, below it, the creation-time values */
        double *M = (double *)malloc(sizeof(double) * (size_t)d * d);
        for (int j = 0; j < d; ++j)
            for (int i = 0; i <= j; ++i)
            {
                const size_t p = (size_t)j * (j + 1) / 2 + i;
                M[(size_t)i * d + j] = f->H[p] + (i == j ? shift : 0.0);
                if (i != j)
                    M[(size_t)j * d + i] = f->H0[p];
            }
        orc_gauss_solve(M, g, d, step);
        free(M);
    }
    if (!f->ls)
    {
        for (int i = 0; i < d; ++i)
            f->x[i] = f->x[i] - step[i];
    }
    else
    {
        /* TrainLoop.h:1550-1637: t = 1, f(x_k) from the round, d^k = -dx;
         * accept the first t with f(x_k - t dx) <= f(x_k) + c t (-(dx, g)),
         * else t *= gamma; after max_steps the last trial point is kept */
        double dot = 0.0;
        for (int i = 0; i < d; ++i)
            dot = dot + step[i] * g[i];
        const double dd = -dot;
        double t = 1.0;
        double *xprev = (double *)malloc(sizeof(double) * d);
        memcpy(xprev, f->x, sizeof(double) * d);
        int kls = 0;
        while (kls < f->ls_max)
        {
            for (int i = 0; i < d; ++i)
                f->x[i] = xprev[i] - t * step[i];
            double freq = 0.0;
            for (int c = 0; c < n; ++c)
            {
                double fi = 0.0;
                orc_client_eval(f->feat + (size_t)c * m * d, f->labels + (size_t)c * m, d, m, f->lambda, f->x,
                                NULL, NULL, NULL, NULL, &fi, NULL);
                freq = freq + fi;
            }
            freq = freq * f->inv_n;
            ++kls;
            if (freq <= fagg + (f->ls_c * t) * dd)
                break;
            t = t * f->ls_gamma;
        }
        f->ls_count = kls;
        free(xprev);
    }

    double gg = 0.0;
    for (int i = 0; i < d; ++i)
        gg = gg + g[i] * g[i];
    if (grad) memcpy(grad, g, sizeof(double) * d);
    if (dx) memcpy(dx, step, sizeof(double) * d);
    if (lk) memcpy(lk, f->lks, sizeof(double) * n);
    if (cnt) memcpy(cnt, f->cnt, sizeof(int) * n);
    for (int c = 0; c < n; ++c)
    {
        if (pos) memcpy(pos + (size_t)c * k, f->pos + (size_t)c * cap, sizeof(uint32_t) * (f->cnt[c] < (int)k ? f->cnt[c] : (int)k));
        if (vals) memcpy(vals + (size_t)c * k, f->vals + (size_t)c * cap, sizeof(double) * (f->cnt[c] < (int)k ? f->cnt[c] : (int)k));
    }
    if (stats)
    {
        stats[0] = sqrt(gg);
        stats[1] = (tracking || f->ls) ? fagg : 0.0;
        stats[2] = xnorm;
    }
    free(g);
    free(U);
    free(step);
    return ok;
}

/* y = L x for the full symmetric matrix held as packed upper L, columns j
 * ascending (MatrixNMD.h:1299-1333 on the full matrix). */
static void sym_matvec_packed(const double *L, int d, const double *x, double *y)
{
    for (int i = 0; i < d; ++i)
        y[i] = 0.0;
    for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i)
        {
            const size_t p = i <= j ? (size_t)j * (j + 1) / 2 + i : (size_t)i * (i + 1) / 2 + j;
            y[i] = y[i] + L[p] * x[j];
        }
}

/* Turns the run into FedNL-PP with tau clients per round and the client
 * sampler seeded `seed` (bin_fednl_local_pp: Args.h:242-264,
 * TrainLoop.h:279-287). Every client starts from l_i = 0 and the direction
 * H_i x - grad f_i(x) at the current x; the master's averages start from
 * their means (LocalWorkerForFedNL_TopK.h:20-30; TrainLoop.h:1060-1080). */
ORC_API void orc_fednl_set_pp(void *h, int tau, uint32_t seed)
{
    orc_fednl *f = (orc_fednl *)h;
    const int n = f->n, m = f->m, d = f->d;
    const size_t S = f->S;
    f->pp_tau = tau;
    f->pp_rng = seed;
    f->pp_dir = (double *)calloc((size_t)n * d, sizeof(double));
    f->pp_lk = (double *)calloc(n, sizeof(double));
    f->pp_g = (double *)calloc(d, sizeof(double));
    f->pp_gmsg = (double *)calloc((size_t)n * d, sizeof(double));
    f->pp_lkmsg = (double *)calloc(n, sizeof(double));
    f->pp_sel = (int *)calloc(n, sizeof(int));
    double *gr = (double *)malloc(sizeof(double) * d);
    double *mv = (double *)malloc(sizeof(double) * d);
    for (int c = 0; c < n; ++c)
    {
        orc_client_eval(f->feat + (size_t)c * m * d, f->labels + (size_t)c * m, d, m, f->lambda, f->x, NULL,
                        NULL, gr, NULL, NULL, NULL);
        sym_matvec_packed(f->learned + (size_t)c * S, d, f->x, mv);
        for (int i = 0; i < d; ++i)
            f->pp_dir[(size_t)c * d + i] = mv[i] - gr[i];
    }
    double li = 0.0;
    for (int c = 0; c < n; ++c)
    {
        for (int i = 0; i < d; ++i)
            f->pp_g[i] = f->pp_g[i] + f->pp_dir[(size_t)c * d + i];
        li = li + f->pp_lk[c];
    }
    for (int i = 0; i < d; ++i)
        f->pp_g[i] = f->pp_g[i] * f->inv_n;
    f->pp_li = li * f->inv_n;
    free(gr);
    free(mv);
}

/* One FedNL-PP round (bin_fednl_local_pp/include/TrainLoop.h:1148-1720,
 * LocalWorkerForFedNL_*.h). Selected clients: the FedNL client step, then
 * (option b) Lk_new = ||learned_new - H_i(x)||_F and the message
 * Lk_new - Lk_old, and the direction l_i x + learned_new x - grad f_i(x),
 * sent as its change. Master, clients in order: g = (sum of changes) / n +
 * g_prev, l = (sum of Lk changes) / n + l_prev, H[idx] += val, then
 * x = (H + shift I)^-1 g (Cholesky, Gauss-Jordan if it fails; shift = l or
 * lambda). Outputs: sel[n] (0/1), cnt[n] (items each selected client sent,
 * 0 for the others), grad[d] = the full gradient at x_r (all
 * clients, sum / n), stats[3] = (||grad||, f or 0, ||x_r||), where f sums
 * the selected clients' f_i / n (TrainLoop.h:1648-1676). Returns 1 if the
 * factorization succeeded. */
ORC_API int orc_fednl_pp_round(void *h, int tracking, int *sel, int *cnt, double *grad, double *stats)
{
    orc_fednl *f = (orc_fednl *)h;
    const int n = f->n, m = f->m, d = f->d;
    const size_t S = f->S, k = (size_t)f->k;
    const size_t cap = S > k ? S : k;
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (int c = 0; c < n; ++c)
        order[c] = (uint32_t)c;
    orc_shuffle_partial(order, (size_t)n, (size_t)f->pp_tau, &f->pp_rng);
    for (int c = 0; c < n; ++c)
        f->pp_sel[c] = 0;
    for (int q = 0; q < f->pp_tau; ++q)
        f->pp_sel[order[q]] = 1;
    free(order);

    double xx = 0.0;
    for (int i = 0; i < d; ++i)
        xx = xx + f->x[i] * f->x[i];
    const double xnorm = sqrt(xx);
    double *tg = (double *)calloc(d, sizeof(double));
    double *gr = (double *)malloc(sizeof(double) * d);
    for (int c = 0; c < n; ++c)
    {
        orc_client_eval(f->feat + (size_t)c * m * d, f->labels + (size_t)c * m, d, m, f->lambda, f->x, NULL,
                        NULL, gr, NULL, NULL, NULL);
        for (int i = 0; i < d; ++i)
            tg[i] = tg[i] + gr[i];
    }
    for (int i = 0; i < d; ++i)
        tg[i] = tg[i] / (double)n;

    double *res = (double *)malloc(sizeof(double) * S);
    double *mv = (double *)malloc(sizeof(double) * d);
    for (int c = 0; c < n; ++c)
    {
        if (!f->pp_sel[c])
            continue;
        client_step(f, c, tracking);
        const double *L = f->learned + (size_t)c * S;
        if (f->option_b)
        {
            for (size_t p = 0; p < S; ++p)
                res[p] = L[p] - f->hx[p];
            const double lk_new = orc_frob_upper_packed(res, d);
            f->pp_lkmsg[c] = lk_new - f->pp_lk[c];
            f->pp_lk[c] = lk_new;
        }
        sym_matvec_packed(L, d, f->x, mv);
        const double *gc = f->grads + (size_t)c * d;
        for (int i = 0; i < d; ++i)
        {
            const double dir = (f->x[i] * f->pp_lk[c] + mv[i]) - gc[i];
            f->pp_gmsg[(size_t)c * d + i] = dir - f->pp_dir[(size_t)c * d + i];
            f->pp_dir[(size_t)c * d + i] = dir;
        }
    }
    free(res);
    free(mv);

    double *g = (double *)calloc(d, sizeof(double));
    double li = 0.0, fagg = 0.0;
    for (int c = 0; c < n; ++c)
    {
        if (!f->pp_sel[c])
            continue;
        for (int i = 0; i < d; ++i)
            g[i] = g[i] + f->pp_gmsg[(size_t)c * d + i];
        const uint32_t *pc = f->pos + (size_t)c * cap;
        const double *vc = f->vals + (size_t)c * cap;
        for (int q = 0; q < f->cnt[c]; ++q)
            f->H[pc[q]] = f->H[pc[q]] + vc[q];
        li = li + f->pp_lkmsg[c];
        fagg = fagg + f->fis[c];
    }
    for (int i = 0; i < d; ++i)
    {
        g[i] = g[i] * f->inv_n;
        g[i] = g[i] + f->pp_g[i];
        f->pp_g[i] = g[i];
    }
    if (f->option_b)
    {
        li = li * f->inv_n;
        li = li + f->pp_li;
        f->pp_li = li;
    }
    fagg = fagg * f->inv_n;

    const double shift = f->option_b ? li : f->lambda;
    double *U = (double *)malloc(sizeof(double) * S);
    int ok = orc_cholesky_packed(f->H, d, shift, U);
    if (ok)
        orc_tri_solve_packed(U, d, g, f->x);
    else
    {
        double *M = (double *)malloc(sizeof(double) * (size_t)d * d);
        for (int j = 0; j < d; ++j)
            for (int i = 0; i <= j; ++i)
            {
                const size_t p = (size_t)j * (j + 1) / 2 + i;
                M[(size_t)i * d + j] = f->H[p] + (i == j ? shift : 0.0);
                if (i != j)
                    M[(size_t)j * d + i] = f->H0[p];
            }
        orc_gauss_solve(M, g, d, f->x);
        free(M);
    }
    free(U);
    double gg = 0.0;
    for (int i = 0; i < d; ++i)
        gg = gg + tg[i] * tg[i];
    if (sel)
        memcpy(sel, f->pp_sel, sizeof(int) * n);
    if (cnt)
        for (int c = 0; c < n; ++c)
            cnt[c] = f->pp_sel[c] ? f->cnt[c] : 0;
    if (grad)
        memcpy(grad, tg, sizeof(double) * d);
    if (stats)
    {
        stats[0] = sqrt(gg);
        stats[1] = tracking ? fagg : 0.0;
        stats[2] = xnorm;
    }
    free(g);
    free(tg);
    free(gr);
    return ok;
}

/* ------------------------------------------------------------------------ */
/* GD (--algorithm gd): LocalWorkerForGD.h:8-49, master TrainLoop.h:993-1052 */
/* ------------------------------------------------------------------------ */
/* rounds of x -= step * (sum_c grad f_c(x)) / n from x = 0, clients summed in
 * order. Per round r: gnorm[r] = ||g||, fval[r] = the aggregate the reference
 * records under tracking ((sum_c (f_c + sizeof(double))) / n,
 * TrainLoop.h:1023-1030; 0 without tracking), xnorm[r] = ||x|| before the
 * step (:1002). x_out: the last iterate. */
ORC_API void orc_gd_run(const double *feat, const double *labels, int n, int m, int d, double lambda,
                        double step, int rounds, int tracking, double *gnorm, double *fval, double *xnorm,
                        double *x_out)
{
    double *x = (double *)calloc((size_t)d, sizeof(double));
    double *g = (double *)malloc(sizeof(double) * d);
    double *gc = (double *)malloc(sizeof(double) * d);
    double inv_n = 1.0 / (double)n;
    for (int r = 0; r < rounds; ++r)
    {
        double xx = 0.0;
        for (int i = 0; i < d; ++i)
            xx = xx + x[i] * x[i];
        for (int i = 0; i < d; ++i)
            g[i] = 0.0;
        double fa = 0.0;
        for (int c = 0; c < n; ++c)
        {
            double fc = 0.0;
            orc_client_eval(feat + (size_t)c * m * d, labels + (size_t)c * m, d, m, lambda, x, NULL, NULL, gc,
                            NULL, &fc, NULL);
            for (int i = 0; i < d; ++i)
                g[i] = g[i] + gc[i];
            if (tracking)
            {
                fa = fa + fc;
                fa = fa + (double)sizeof(double);
            }
        }
        double gg = 0.0;
        for (int i = 0; i < d; ++i)
        {
            g[i] = g[i] * inv_n;
            gg = gg + g[i] * g[i];
        }
        gnorm[r] = sqrt(gg);
        fval[r] = tracking ? fa * inv_n : 0.0;
        xnorm[r] = sqrt(xx);
        /* subInPlaceVectorWithMultiple: x_i - step * g_i */
        for (int i = 0; i < d; ++i)
            x[i] = x[i] - step * g[i];
    }
    if (x_out)
        memcpy(x_out, x, sizeof(double) * d);
    free(x);
    free(g);
    free(gc);
}
