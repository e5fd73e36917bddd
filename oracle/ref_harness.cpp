// Checker-only harness around the reference's own FedNL building blocks.
//
// This file is test infrastructure. It is compiled by oracle/Makefile.ref
// against the reference headers where they lie (dopt/dopt/...), into
// oracle/_ref/libdopt_ref.so, and is loaded only by tests/, by the golden
// fixture script and by bench.py's CPU-baseline leg. Nothing in the product
// path links or loads it.
//
// The reference binary bin_fednl_local itself cannot be built here without
// its CMake-generated Version.h (dopt/bin_fednl_local/include/Reports.h), so
// the round orchestration below restates the FedNL branch of
//   dopt/bin_fednl_local/include/TrainLoop.h:1054-1698
// with `--debug forcesequpdate` semantics (master consumes clients 0..n-1 in
// order, server worker pools off), while every numeric kernel (logistic
// oracle, Hessian, compressors, Cholesky, triangular solves, RNGs) is the
// reference's own template code.
#include "dopt/linalg_vectors/include/VectorND_Raw.h"
#include "dopt/linalg_matrices/include/MatrixNMD.h"
#include "dopt/optimization_problems/include/ml/logistic_regression.h"
#include "dopt/math_routines/include/SpecialMathRoutinesForMatrix.h"
#include "dopt/math_routines/include/MinHeap.h"
#include "dopt/random/include/RandomGenIntegerLinear.h"
#include "dopt/random/include/RandomGenRealLinear.h"
#include "dopt/random/include/Shuffle.h"
#include "dopt/linalg_matrices/include/factorization/CholeskyFactorization.h"
#include "dopt/linalg_linsolvers/include/ElementarySolvers.h"
#include "dopt/linalg_linsolvers/include/GaussEliminationSolvers.h"
#include "dopt/optimization_problems/include/ml/dataset_loader_fs.h"
#include "dopt/fs/include/FileSystemHelpers.h"
#include "dopt/system/include/digest/Crc.h"

#include <stdint.h>
#include <string.h>
#include <math.h>

#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

typedef dopt::VectorNDRaw_d Vec;
typedef dopt::MatrixNMD<Vec> Mat;
typedef dopt::L2RegulirizeLogisticRegression<Mat, Vec, double> Problem;

namespace
{
    // One OS thread that runs submitted jobs synchronously. The TopLEK
    // selector keeps a thread_local generator (SpecialMathRoutinesForMatrix.h:383),
    // so a reference "worker" must be one persistent thread.
    class Worker
    {
    public:
        Worker() : th_([this] { loop(); }) {}
        ~Worker()
        {
            {
                std::lock_guard<std::mutex> g(m_);
                stop_ = true;
            }
            cv_.notify_all();
            th_.join();
        }
        void submit(std::function<void()> f)
        {
            {
                std::lock_guard<std::mutex> g(m_);
                job_ = std::move(f);
                has_ = true;
                done_ = false;
            }
            cv_.notify_all();
        }
        void wait()
        {
            std::unique_lock<std::mutex> g(m_);
            cv_.wait(g, [this] { return done_; });
        }

    private:
        void loop()
        {
            for (;;)
            {
                std::function<void()> f;
                {
                    std::unique_lock<std::mutex> g(m_);
                    cv_.wait(g, [this] { return has_ || stop_; });
                    if (stop_ && !has_)
                        return;
                    f = std::move(job_);
                    has_ = false;
                }
                f();
                {
                    std::lock_guard<std::mutex> g(m_);
                    done_ = true;
                }
                cv_.notify_all();
            }
        }
        std::mutex m_;
        std::condition_variable cv_;
        std::function<void()> job_;
        bool has_ = false;
        bool done_ = true;
        bool stop_ = false;
        std::thread th_;
    };

    Mat matFromDense(const double* a, int d)
    {
        // a is d x d column-major with leading dimension lda_ref(d)
        Mat m(d, d);
        memcpy(m.matrixByCols.data(), a, sizeof(double) * m.LDA * d);
        return m;
    }

    // features: m samples x d, row-major (sample s at a + s*d); builds the
    // d x m transposed design matrix the reference stores (dataset_loader_fs.h).
    Mat designTr(const double* a, int d, int m)
    {
        Mat t(d, m);
        for (int s = 0; s < m; ++s)
            for (int i = 0; i < d; ++i)
                t.set(i, s, a[(size_t)s * d + i]);
        return t;
    }

    Vec vecFrom(const double* v, int n)
    {
        Vec r(n);
        for (int i = 0; i < n; ++i)
            r[i] = v[i];
        return r;
    }
}

extern "C"
{
    // dopt's own LIBSVM loading as train() runs it (TrainLoop.h:123-170): the
    // same separators, the statistics pass, then createDataset with the
    // intercept flag and the binary label remapping. X (N x d row-major:
    // sample, attribute) and y; X == NULL: sizes only. crc_x / crc_y (may be
    // NULL) get the digests Reports.h:25-33 prints (matrixByCols raw data,
    // LDA padding included, and the labels). Returns 0, -3 (no file), -4.
    int ref_load_libsvm(const char* path, int add_intercept, int* N, int* d, double* X, double* y,
                        uint32_t* crc_x, uint32_t* crc_y)
    {
        auto line_sep = [](char c) { return c == '\n' || c == '\r'; };
        auto field_sep = [](char c) { return c == '\t' || c == ' '; };
        auto kv_sep = [](char c) { return c == ':'; };
        auto skip = [](char c) { return false; };
        dopt::FileSystemHelpers::FileMappingResult m = dopt::FileSystemHelpers::mapFileToMemory(path, true);
        if (!m.isOk)
            return -3;
        dopt::Data raw(m.memory, m.memorySizeInBytes, dopt::Data::MemInitializedType::eGiftWholeMemoryPleaseNotFree);
        dopt::TextFileDataSetLoader loader(raw, line_sep, field_sep, kv_sep, skip);
        auto stats = loader.computeStatisticsWithOneDataPass<false, true, int32_t, int32_t>();
        if (!stats.ok) {
            raw.clear();
            dopt::FileSystemHelpers::unmapFileFromMemory(m);
            return -4;
        }
        typedef dopt::DataSet<Mat, Vec, dopt::VectorNDRaw_b, Vec> DS;
        DS ds;
        const int flags = (add_intercept ? dopt::CreateDatasetFlags::eAddInterceptTerm : dopt::CreateDatasetFlags::eNone) |
                          dopt::CreateDatasetFlags::eMakeRemappingForBinaryLogisticLoss;
        loader.createDataset<true, true>(ds, stats, flags);
        raw.clear();
        dopt::FileSystemHelpers::unmapFileFromMemory(m);
        const size_t n = ds.train_samples_tr.columns(), dd = ds.train_samples_tr.rows();
        *N = (int)n;
        *d = (int)dd;
        if (crc_x)
            *crc_x = dopt::crc32(ds.train_samples_tr.matrixByCols.rawData(),
                                 ds.train_samples_tr.matrixByCols.sizeInBytes(), dopt::crc32Seed());
        if (crc_y)
            *crc_y = dopt::crc32(ds.train_outputs.rawData(), ds.train_outputs.sizeInBytes(), dopt::crc32Seed());
        if (X) {
            for (size_t s_ = 0; s_ < n; ++s_) {
                for (size_t a = 0; a < dd; ++a)
                    X[s_ * dd + a] = ds.train_samples_tr.get(a, s_);
                y[s_] = ds.train_outputs[s_];
            }
        }
        return 0;
    }

    // dopt::crc32 / crc32Seed (Crc.cpp:74-90)
    uint32_t ref_crc32(const void* buf, size_t size, uint32_t seed) { return dopt::crc32(buf, size, seed); }
    uint32_t ref_crc32_seed() { return dopt::crc32Seed(); }

    int ref_lda(int d)
    {
        Mat m(d, d);
        return (int)m.LDA;
    }

    void ref_upper_indices(int d, uint32_t* out)
    {
        Mat m(d, d);
        std::vector<uint32_t> v = dopt::indiciesForUpperTriangularPart(m);
        memcpy(out, v.data(), v.size() * sizeof(uint32_t));
    }

    // TopK with ignoreSign=true on a full d x LDA matrix (SpecialMathRoutinesForMatrix.h:229-366)
    int ref_topk(const double* mat, int d, int k, uint32_t* out)
    {
        Mat m = matFromDense(mat, d);
        std::vector<uint32_t> up = dopt::indiciesForUpperTriangularPart(m);
        std::vector<uint32_t> r = dopt::getTopKFromUpperDiagonalPart<true>(m, (size_t)k, up);
        memcpy(out, r.data(), r.size() * sizeof(uint32_t));
        return (int)r.size();
    }

    // TopK with ignoreSign=false (the signed variant of the same routine)
    int ref_topk_signed(const double* mat, int d, int k, uint32_t* out)
    {
        Mat m = matFromDense(mat, d);
        std::vector<uint32_t> up = dopt::indiciesForUpperTriangularPart(m);
        std::vector<uint32_t> r = dopt::getTopKFromUpperDiagonalPart<false>(m, (size_t)k, up);
        memcpy(out, r.data(), r.size() * sizeof(uint32_t));
        return (int)r.size();
    }

    // A stream of TopLEK calls on one fresh worker thread, so the thread_local
    // RealLinear generator starts from its default seed 1234 and advances
    // across calls exactly as in one reference worker.
    void ref_toplek_stream(int ncalls, const double* mats, int d, int k, double delta,
                           uint32_t* out_idx, int* out_cnt)
    {
        Worker w;
        w.submit([&] {
            Mat probe(d, d);
            size_t stride = probe.LDA * (size_t)d;
            std::vector<uint32_t> up = dopt::indiciesForUpperTriangularPart(probe);
            for (int c = 0; c < ncalls; ++c)
            {
                Mat m = matFromDense(mats + stride * c, d);
                std::vector<uint32_t> r = dopt::getTopLEKFromUpperDiagonalPart<true>(m, (size_t)k, up, delta);
                out_cnt[c] = (int)r.size();
                if (!r.empty())
                    memcpy(out_idx + (size_t)k * c, r.data(), r.size() * sizeof(uint32_t));
            }
        });
        w.wait();
    }

    // RandK index stream of one client: `rounds` calls of
    // generateRandKItemsInUpperTriangularPart with one RandomGenIntegerLinear seeded `seed`.
    void ref_randk(uint32_t seed, int rounds, int d, int k, uint32_t* out)
    {
        Mat m(d, d);
        std::vector<uint32_t> up = dopt::indiciesForUpperTriangularPart(m);
        dopt::RandomGenIntegerLinear g;
        g.setSeed(seed);
        for (int r = 0; r < rounds; ++r)
        {
            std::vector<uint32_t> v = dopt::generateRandKItemsInUpperTriangularPart(g, (size_t)k, m, up);
            memcpy(out + (size_t)r * k, v.data(), (size_t)k * sizeof(uint32_t));
        }
    }

    void ref_randseqk(uint32_t seed, int rounds, int d, int k, uint32_t* out)
    {
        Mat m(d, d);
        std::vector<uint32_t> up = dopt::indiciesForUpperTriangularPart(m);
        dopt::RandomGenIntegerLinear g;
        g.setSeed(seed);
        for (int r = 0; r < rounds; ++r)
        {
            std::vector<uint32_t> v = dopt::generateRandSeqKItemsInUpperTriangularPart(g, (size_t)k, m, up);
            memcpy(out + (size_t)r * k, v.data(), (size_t)k * sizeof(uint32_t));
        }
    }

    void ref_lcg_integer(uint32_t seed, int n, uint32_t* out_int, double* out_real)
    {
        dopt::RandomGenIntegerLinear g;
        g.setSeed(seed);
        dopt::RandomGenIntegerLinear h;
        h.setSeed(seed);
        for (int i = 0; i < n; ++i)
        {
            out_int[i] = g.generateInteger();
            out_real[i] = h.generateRealInUnitInterval<double>();
        }
    }

    void ref_lcg_real(uint32_t seed, int n, double* out)
    {
        dopt::RandomGenRealLinear g;
        g.setSeed(seed);
        for (int i = 0; i < n; ++i)
            out[i] = g.generateRealInUnitInterval();
    }

    // Full Fisher-Yates permutation of 0..n-1 with partial stop k (Shuffle.h:31-93).
    void ref_shuffle(uint32_t seed, int n, int k, uint32_t* out)
    {
        std::vector<uint32_t> v(n);
        for (int i = 0; i < n; ++i)
            v[i] = (uint32_t)i;
        dopt::RandomGenIntegerLinear g;
        g.setSeed(seed);
        dopt::shuffle(v, (size_t)k, g);
        memcpy(out, v.data(), sizeof(uint32_t) * n);
    }

    // L(f) of logistic_regression.h:707-740 on m samples (features m x d
    // row-major, labels +-1) with the tolerance train() uses, 1e-3 mu
    // (TrainLoop.h:306-316); *iters gets the power iterations.
    double ref_l_smoothness(const double* feat, const double* labels, int d, int m, double lambda, int* iters)
    {
        Mat atr = designTr(feat, d, m);
        Vec b = vecFrom(labels, m);
        Problem p(atr, b, lambda);
        size_t it = 0;
        const double L = p.computeLSmoothness(0.001 * p.computeMuStrongConvexity(), &it);
        *iters = (int)it;
        return L;
    }

    // Logistic oracle of one client at x (logistic_regression.h:88-547).
    // hess is d x LDA (full symmetric, +lambda on the diagonal).
    void ref_client_eval(const double* feat, const double* labels, int d, int m, double lambda,
                         const double* x, double* margin, double* sigm, double* grad,
                         double* hess, double* fval)
    {
        Mat atr = designTr(feat, d, m);
        Vec b = vecFrom(labels, m);
        Problem p(atr, b, lambda);
        Vec xv = vecFrom(x, d);
        Vec t = p.evaluateClassificationMargin(xv);
        Vec s = p.evaluateClassificationMarginSigmoid(t);
        Vec g = p.evaluateGradient(xv, t, s);
        Mat h = p.evaluateHessian(xv, t, s);
        for (int i = 0; i < m; ++i)
        {
            margin[i] = t[i];
            sigm[i] = s[i];
        }
        for (int i = 0; i < d; ++i)
            grad[i] = g[i];
        memcpy(hess, h.matrixByCols.dataConst(), sizeof(double) * h.LDA * d);
        *fval = p.evaluateFunction(xv, t, s);
    }

    double ref_frob_upper(const double* mat, int d)
    {
        Mat m = matFromDense(mat, d);
        return m.frobeniusNormForSymmetricMatrixFromUpPart();
    }

    double ref_frob_upper_sqr(const double* mat, int d)
    {
        Mat m = matFromDense(mat, d);
        return m.frobeniusNormSquareForSymmetricMatrixFromUpPart();
    }

    // Cholesky of (H_upper + shift*I) then the two substitutions used by the
    // master (TrainLoop.h:1449-1466). Returns 1 on success, 0 if the
    // factorization reports failure.
    int ref_newton_solve(const double* H, int d, double shift, const double* b, double* dx, double* ltr)
    {
        Mat h = matFromDense(H, d);
        Mat l;
        bool ok = dopt::cholFactorization::choleskyFactorization<Mat, true, true>(l, h, shift);
        if (!ok)
            return 0;
        Vec bv = vecFrom(b, d);
        Vec y = dopt::forwardSubstitutionWithATranspose(l, bv);
        Vec r = dopt::backwardSubstitutionWithATranspose(l, y);
        for (int i = 0; i < d; ++i)
            dx[i] = r[i];
        if (ltr)
            memcpy(ltr, l.matrixByCols.dataConst(), sizeof(double) * l.LDA * d);
        return 1;
    }

    // GD (--algorithm gd) with the reference's own client oracle and vector
    // arithmetic: the workers' evaluateGradient (LocalWorkerForGD.h:30-37) and
    // the master's round (TrainLoop.h:993-1052), clients consumed in order.
    // Per round: ||g||, the recorded f aggregate (each client's f_c and then
    // sizeof(double) added, divided by n; 0 without tracking), ||x|| before
    // the step. x_out: the last iterate.
    void ref_gd_run(const double* feat, const double* labels, int n, int m, int d, double lambda, double step,
                    int rounds, int tracking, double* grad_norm, double* fval, double* x_norm, double* x_out)
    {
        std::vector<Problem*> probs(n);
        for (int c = 0; c < n; ++c)
        {
            Mat atr = designTr(feat + (size_t)c * m * d, d, m);
            Vec b = vecFrom(labels + (size_t)c * m, m);
            probs[c] = new Problem(atr, b, lambda);
        }
        Vec xCur(d);
        Vec g(d);
        const double inv_n = 1.0 / double(n);
        for (int r = 0; r < rounds; ++r)
        {
            const double xnorm = xCur.vectorL2Norm();
            g.setAllToDefault();
            double fa = 0.0;
            for (int c = 0; c < n; ++c)
            {
                g += probs[c]->evaluateGradient(xCur);
                if (tracking)
                {
                    fa += probs[c]->evaluateFunction(xCur);
                    fa += sizeof(fa);
                }
            }
            g *= inv_n;
            if (tracking)
                fa *= inv_n;
            grad_norm[r] = g.vectorL2Norm();
            fval[r] = fa;
            x_norm[r] = xnorm;
            xCur.subInPlaceVectorWithMultiple(step, g);
        }
        for (int i = 0; i < d; ++i)
            x_out[i] = xCur[i];
        for (int c = 0; c < n; ++c)
            delete probs[c];
    }

    // compressor codes follow UsedTypes.h:17-25
    enum { C_IDENTICAL = 0, C_RANDK = 1, C_TOPK = 2, C_RANDSEQK = 3, C_TOPLEK = 4, C_NATURAL = 5 };

    // FedNL (option b when option_b != 0, else option a) with
    // forcesequpdate semantics. Clients are assigned round-robin to
    // `workers` threads (TrainLoop.h:664-686) which compute their clients in
    // ascending order; the master then consumes clients 0..n-1 in order.
    //
    // feat: n*m*d (client-major, then sample-major), labels: n*m.
    // Per-round outputs (may be NULL): grad_norm[r], fval[r], x_norm[r].
    // Final iterate in x_out (d). Lockstep dumps for the first `dump_rounds`
    // rounds (may be 0): dump_grad[r*d], dump_lk[r*n+c], dump_cnt[r*n+c],
    // dump_idx[(r*n+c)*k], dump_dx[r*d].
    // Returns 0, or -7 on an internal error, or 1 if a factorization failed.
    int ref_fednl_run(const double* feat, const double* labels, int n, int m, int d,
                      double lambda, int compressor, int k, double alpha, int option_b,
                      int rounds, int workers, int tracking,
                      double* grad_norm, double* fval, double* x_norm, double* x_out,
                      int dump_rounds, double* dump_grad, double* dump_lk, int* dump_cnt,
                      uint32_t* dump_idx, double* dump_dx, double* seconds,
                      int ls_max, double ls_c, double ls_gamma, int* ls_counts, int* chol_ok)
    {
        // ls_max > 0: FedNL-LS (TrainLoop.h:344-355, 1546-1640)
        if (workers < 1)
            workers = 1;
        if (workers > n)
            workers = n;

        std::vector<Problem*> probs(n);
        std::vector<Mat> learned(n);
        Mat H = Mat::getZeroSquareMatrix(d);
        Vec x0(d);
        for (int c = 0; c < n; ++c)
        {
            Mat atr = designTr(feat + (size_t)c * m * d, d, m);
            Vec b = vecFrom(labels + (size_t)c * m, m);
            probs[c] = new Problem(atr, b, lambda);
            learned[c] = probs[c]->evaluateHessian(x0);
            H += learned[c];
        }
        double inv_n = 1.0 / double(n);
        H *= inv_n;
        double prescale = inv_n * alpha;
        size_t S = ((size_t)(d + 1) * d) / 2;
        double mult = double(S) / double(k);
        double delta_toplek = double(k) * (1.0 / double(S));

        std::vector<uint32_t> up = dopt::indiciesForUpperTriangularPart(H);
        std::vector<dopt::RandomGenIntegerLinear> gens(n);
        for (int c = 0; c < n; ++c)
            gens[c].setSeed((uint32_t)(123 + c));

        Vec x = x0;
        std::vector<Vec> grads(n);
        std::vector<double> lks(n), fis(n);
        std::vector<std::vector<uint32_t>> idxs(n);
        std::vector<std::vector<double>> vals(n);

        std::vector<Worker*> pool(workers);
        for (int w = 0; w < workers; ++w)
            pool[w] = new Worker();

        auto t0 = std::chrono::steady_clock::now();
        int status = 0;

        for (int r = 0; r < rounds; ++r)
        {
            double xnorm = x.vectorL2Norm();
            for (int w = 0; w < workers; ++w)
            {
                pool[w]->submit([&, w] {
                    for (int c = w; c < n; c += workers)
                    {
                        Problem* p = probs[c];
                        Vec t = p->evaluateClassificationMargin(x);
                        Vec s = p->evaluateClassificationMarginSigmoid(t);
                        grads[c] = p->evaluateGradient(x, t, s);
                        Mat hx = p->evaluateHessian(x, t, s);
                        Mat diff = hx - learned[c];
                        lks[c] = diff.frobeniusNormForSymmetricMatrixFromUpPart();
                        std::vector<uint32_t> kc;
                        double scale = 1.0;
                        if (compressor == C_TOPK)
                            kc = dopt::getTopKFromUpperDiagonalPart<true>(diff, (size_t)k, up);
                        else if (compressor == C_TOPLEK)
                            kc = dopt::getTopLEKFromUpperDiagonalPart<true>(diff, (size_t)k, up, delta_toplek);
                        else if (compressor == C_RANDK)
                        {
                            kc = dopt::generateRandKItemsInUpperTriangularPart(gens[c], (size_t)k, learned[c], up);
                            scale = mult;
                        }
                        else if (compressor == C_RANDSEQK)
                        {
                            kc = dopt::generateRandSeqKItemsInUpperTriangularPart(gens[c], (size_t)k, learned[c], up);
                            scale = mult;
                        }
                        else
                            kc = up;
                        // Natural: the reference's own masking of the whole difference
                        // (LocalWorkerForFedNL_Natural.h:73-75)
                        if (compressor == C_NATURAL)
                            diff.applyNaturalCompressor();
                        std::vector<double> vc(kc.size());
                        for (size_t q = 0; q < kc.size(); ++q)
                        {
                            double item = diff.matrixByCols.get(kc[q]);
                            if (scale != 1.0)
                                item = item * scale;
                            vc[q] = item * prescale;
                        }
                        for (size_t q = 0; q < kc.size(); ++q)
                        {
                            size_t first = kc[q];
                            size_t iRow = 0, iCol = 0;
                            diff.getPositionFromFlatternIndex(iRow, iCol, first);
                            double item = diff.matrixByCols.get(first);
                            if (scale != 1.0)
                                item = item * scale;
                            if (iRow == iCol)
                                learned[c].matrixByCols[first] += alpha * item;
                            else
                            {
                                learned[c].matrixByCols[first] += alpha * item;
                                size_t second = diff.getFlattenIndexFromPosition(iCol, iRow);
                                learned[c].matrixByCols[second] += alpha * item;
                            }
                        }
                        idxs[c].swap(kc);
                        vals[c].swap(vc);
                        if (tracking || ls_max > 0)
                            fis[c] = p->evaluateFunction(x, t, s);
                    }
                });
            }
            for (int w = 0; w < workers; ++w)
                pool[w]->wait();

            // master, clients in order 0..n-1 (TrainLoop.h:1094-1333)
            Vec g(d);
            double lsum = 0.0;
            for (int c = 0; c < n; ++c)
            {
                for (int i = 0; i < d; ++i)
                    g[i] += grads[c][i];
                for (size_t q = 0; q < idxs[c].size(); ++q)
                    H.matrixByCols[idxs[c][q]] += vals[c][q];
                lsum += lks[c];
            }
            for (int i = 0; i < d; ++i)
                g[i] *= inv_n;
            lsum *= inv_n;

            double shift = option_b ? lsum : probs[0]->computeMuStrongConvexity();
            Mat ltr;
            bool ok = dopt::cholFactorization::choleskyFactorization<Mat, true, true>(ltr, H, shift);
            Vec dx(d);
            if (ok)
            {
                Vec y = dopt::forwardSubstitutionWithATranspose(ltr, g);
                dx = dopt::backwardSubstitutionWithATranspose(ltr, y);
            }
            else
            {
                // TrainLoop.h:1410-1420, 1476-1486: Gauss elimination on the
                // master's full H (its lower triangle never updated) + shift I
                Mat slau = H;
                slau.addToAllDiagonalEntries(shift);
                dx = dopt::gausEleminationSolver(slau, g);
                status = 1;
            }
            if (chol_ok)
                chol_ok[r] = ok ? 1 : 0;

            if (r < dump_rounds)
            {
                for (int i = 0; i < d; ++i)
                {
                    dump_grad[(size_t)r * d + i] = g[i];
                    dump_dx[(size_t)r * d + i] = dx[i];
                }
                for (int c = 0; c < n; ++c)
                {
                    dump_lk[(size_t)r * n + c] = lks[c];
                    dump_cnt[(size_t)r * n + c] = (int)idxs[c].size();
                    if (!idxs[c].empty())
                        memcpy(dump_idx + ((size_t)r * n + c) * k, idxs[c].data(), idxs[c].size() * sizeof(uint32_t));
                }
            }

            double fagg = 0.0;
            if (tracking || ls_max > 0)
            {
                for (int c = 0; c < n; ++c)
                    fagg += fis[c];
                fagg *= inv_n;
            }
            if (ls_max <= 0)
            {
                x -= dx;
            }
            else
            {
                // TrainLoop.h:1550-1637, with the reference's own vector ops
                double t = 1.0;
                const double dd = -(dx & g);
                Vec xprev = x;
                int kls = 0;
                while (kls < ls_max)
                {
                    x = xprev - (t * dx);
                    double freq = 0.0;
                    for (int c = 0; c < n; ++c)
                    {
                        Problem* p = probs[c];
                        Vec tm = p->evaluateClassificationMargin(x);
                        Vec sm = p->evaluateClassificationMarginSigmoid(tm);
                        freq += p->evaluateFunction(x, tm, sm);
                    }
                    freq *= inv_n;
                    ++kls;
                    if (freq <= fagg + (ls_c * t) * dd)
                        break;
                    t *= ls_gamma;
                }
                if (ls_counts)
                    ls_counts[r] = kls;
            }
            if (grad_norm)
                grad_norm[r] = g.vectorL2Norm();
            if (fval)
                fval[r] = fagg;
            if (x_norm)
                x_norm[r] = xnorm;
            if (seconds)
                seconds[r] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }

        for (int i = 0; i < d; ++i)
            x_out[i] = x[i];
        // the reference's vector memory pools are per thread: release what a
        // worker allocated on that worker (as LocalWorkerForFedNL_*.h do on exit)
        for (int w = 0; w < workers; ++w)
            pool[w]->submit([&, w] {
                for (int c = w; c < n; c += workers)
                    grads[c] = Vec();
            });
        for (int w = 0; w < workers; ++w)
            pool[w]->wait();
        for (int w = 0; w < workers; ++w)
            delete pool[w];
        for (int c = 0; c < n; ++c)
            delete probs[c];
        return status;
    }

    // FedNL-PP (dopt/bin_fednl_local_pp): each round `tau` clients drawn by a
    // partial shuffle of 0..n-1 with one RandomGenIntegerLinear seeded
    // `client_seed` (TrainLoop.h:279-287,1155-1157) do the FedNL client step,
    // then send Lk_new - Lk_old (option b) and the change of their direction
    // l_i x + H_i x - grad f_i(x) (LocalWorkerForFedNL_TopK.h:20-30,129-145).
    // The master keeps running averages of both and solves
    // (H + shift I) x_{k+1} = g (TrainLoop.h:1060-1080,1175-1640). Clients of a
    // worker run in ascending order and the master consumes 0..n-1 in order
    // (forcesequpdate). Per round: ||grad f(x_r)|| of all clients, f (sum
    // over the selected clients / n, with tracking), ||x_r||, factorization
    // flag, and the selected clients (sel_out[r*n + c] = 0/1).
    int ref_fednl_pp_run(const double* feat, const double* labels, int n, int m, int d,
                         double lambda, int compressor, int k, double alpha, int option_b,
                         int rounds, int workers, int tracking, int tau, uint32_t client_seed,
                         double* grad_norm, double* fval, double* x_norm, double* x_out, int* chol_ok,
                         int* sel_out)
    {
        if (workers < 1)
            workers = 1;
        if (workers > n)
            workers = n;
        if (tau < 1 || tau > n)
            return -2;

        std::vector<Problem*> probs(n);
        std::vector<Mat> learned(n);
        std::vector<Vec> local_dir(n);
        std::vector<double> local_lk(n, 0.0);
        Mat H = Mat::getZeroSquareMatrix(d);
        Vec x0(d);
        Vec prev_g(d);
        double prev_li = 0.0;
        for (int c = 0; c < n; ++c)
        {
            Mat atr = designTr(feat + (size_t)c * m * d, d, m);
            Vec b = vecFrom(labels + (size_t)c * m, m);
            probs[c] = new Problem(atr, b, lambda);
            learned[c] = probs[c]->evaluateHessian(x0);
        }
        // vectors a worker reassigns are allocated (and at the end released)
        // on that worker: the reference's vector memory pools are per thread
        std::vector<Worker*> pool(workers);
        for (int w = 0; w < workers; ++w)
            pool[w] = new Worker();
        for (int w = 0; w < workers; ++w)
            pool[w]->submit([&, w] {
                for (int c = w; c < n; c += workers)
                    local_dir[c] = Mat::matrixVectorMultiply(learned[c], x0) - probs[c]->evaluateGradient(x0);
            });
        for (int w = 0; w < workers; ++w)
            pool[w]->wait();
        for (int c = 0; c < n; ++c)
        {
            H += learned[c];
            prev_g += local_dir[c];
            prev_li += local_lk[c];
        }
        double inv_n = 1.0 / double(n);
        H *= inv_n;
        prev_g *= inv_n;
        prev_li *= inv_n;
        double prescale = inv_n * alpha;
        size_t S = ((size_t)(d + 1) * d) / 2;
        double mult = double(S) / double(k);
        double delta_toplek = double(k) * (1.0 / double(S));
        std::vector<uint32_t> up = dopt::indiciesForUpperTriangularPart(H);
        std::vector<dopt::RandomGenIntegerLinear> gens(n);
        for (int c = 0; c < n; ++c)
            gens[c].setSeed((uint32_t)(123 + c));
        dopt::RandomGenIntegerLinear sampler;
        sampler.setSeed(client_seed);

        Vec x = x0;
        std::vector<Vec> gmsg(n);
        std::vector<double> lkmsg(n), fis(n);
        std::vector<std::vector<uint32_t>> idxs(n);
        std::vector<std::vector<double>> vals(n);
        std::vector<char> selected(n, 0);
        int status = 0;

        for (int r = 0; r < rounds; ++r)
        {
            std::vector<uint32_t> order(n);
            for (int c = 0; c < n; ++c)
                order[c] = (uint32_t)c;
            dopt::shuffle(order, (size_t)tau, sampler);
            std::fill(selected.begin(), selected.end(), 0);
            for (int q = 0; q < tau; ++q)
                selected[order[q]] = 1;
            if (sel_out)
                for (int c = 0; c < n; ++c)
                    sel_out[(size_t)r * n + c] = selected[c];

            double xnorm = x.vectorL2Norm();
            Vec tg(d);
            for (int c = 0; c < n; ++c)
                tg += probs[c]->evaluateGradient(x);
            tg = tg / double(n);

            for (int w = 0; w < workers; ++w)
            {
                pool[w]->submit([&, w] {
                    for (int c = w; c < n; c += workers)
                    {
                        if (!selected[c])
                            continue;
                        Problem* p = probs[c];
                        Vec t = p->evaluateClassificationMargin(x);
                        Vec s = p->evaluateClassificationMarginSigmoid(t);
                        Mat hx = p->evaluateHessian(x, t, s);
                        Mat diff = hx - learned[c];
                        std::vector<uint32_t> kc;
                        double scale = 1.0;
                        if (compressor == C_TOPK)
                            kc = dopt::getTopKFromUpperDiagonalPart<true>(diff, (size_t)k, up);
                        else if (compressor == C_TOPLEK)
                            kc = dopt::getTopLEKFromUpperDiagonalPart<true>(diff, (size_t)k, up, delta_toplek);
                        else if (compressor == C_RANDK)
                        {
                            kc = dopt::generateRandKItemsInUpperTriangularPart(gens[c], (size_t)k, learned[c], up);
                            scale = mult;
                        }
                        else if (compressor == C_RANDSEQK)
                        {
                            kc = dopt::generateRandSeqKItemsInUpperTriangularPart(gens[c], (size_t)k, learned[c], up);
                            scale = mult;
                        }
                        else
                            kc = up;
                        if (compressor == C_NATURAL)
                            diff.applyNaturalCompressor();
                        std::vector<double> vc(kc.size());
                        for (size_t q = 0; q < kc.size(); ++q)
                        {
                            double item = diff.matrixByCols.get(kc[q]);
                            if (scale != 1.0)
                                item = item * scale;
                            vc[q] = item * prescale;
                        }
                        for (size_t q = 0; q < kc.size(); ++q)
                        {
                            size_t first = kc[q];
                            size_t iRow = 0, iCol = 0;
                            diff.getPositionFromFlatternIndex(iRow, iCol, first);
                            double item = diff.matrixByCols.get(first);
                            if (scale != 1.0)
                                item = item * scale;
                            learned[c].matrixByCols[first] += alpha * item;
                            if (iRow != iCol)
                            {
                                size_t second = diff.getFlattenIndexFromPosition(iCol, iRow);
                                learned[c].matrixByCols[second] += alpha * item;
                            }
                        }
                        if (option_b)
                        {
                            double newLk = (learned[c] - hx).frobeniusNormForSymmetricMatrixFromUpPart();
                            lkmsg[c] = newLk - local_lk[c];
                            local_lk[c] = newLk;
                        }
                        Vec dir = (x * local_lk[c]) + Mat::matrixVectorMultiply(learned[c], x) -
                                  p->evaluateGradient(x, t, s);
                        gmsg[c] = dir - local_dir[c];
                        local_dir[c] = dir;
                        idxs[c].swap(kc);
                        vals[c].swap(vc);
                        if (tracking)
                            fis[c] = p->evaluateFunction(x, t, s);
                    }
                });
            }
            for (int w = 0; w < workers; ++w)
                pool[w]->wait();

            Vec g(d);
            double li = 0.0, fagg = 0.0;
            for (int c = 0; c < n; ++c)
            {
                if (!selected[c])
                    continue;
                g += gmsg[c];
                for (size_t q = 0; q < idxs[c].size(); ++q)
                    H.matrixByCols[idxs[c][q]] += vals[c][q];
                if (option_b)
                    li += lkmsg[c];
                if (tracking)
                    fagg += fis[c];
            }
            g *= inv_n;
            g += prev_g;
            prev_g = g;
            if (option_b)
            {
                li *= inv_n;
                li += prev_li;
                prev_li = li;
            }
            fagg *= inv_n;

            double shift = option_b ? li : probs[0]->computeMuStrongConvexity();
            Mat ltr;
            bool ok = dopt::cholFactorization::choleskyFactorization<Mat, true, true>(ltr, H, shift);
            if (ok)
            {
                Vec y = dopt::forwardSubstitutionWithATranspose(ltr, g);
                x = dopt::backwardSubstitutionWithATranspose(ltr, y);
            }
            else
            {
                Mat slau = H;
                slau.addToAllDiagonalEntries(shift);
                x = dopt::gausEleminationSolver(slau, g);
                status = 1;
            }
            if (chol_ok)
                chol_ok[r] = ok ? 1 : 0;
            if (grad_norm)
                grad_norm[r] = tg.vectorL2Norm();
            if (fval)
                fval[r] = tracking ? fagg : 0.0;
            if (x_norm)
                x_norm[r] = xnorm;
        }

        for (int i = 0; i < d; ++i)
            x_out[i] = x[i];
        for (int w = 0; w < workers; ++w)
            pool[w]->submit([&, w] {
                for (int c = w; c < n; c += workers)
                {
                    gmsg[c] = Vec();
                    local_dir[c] = Vec();
                }
            });
        for (int w = 0; w < workers; ++w)
            pool[w]->wait();
        for (int w = 0; w < workers; ++w)
            delete pool[w];
        for (int c = 0; c < n; ++c)
            delete probs[c];
        return status;
    }
}
