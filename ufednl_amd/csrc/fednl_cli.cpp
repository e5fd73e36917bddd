// fednl_cli.cpp -- the bin_fednl_local command-line surface on top of the
// engine's C ABI: flag parsing (dopt/bin_fednl_local/include/Args.h), LIBSVM
// loading (dopt/dopt/optimization_problems/include/ml/dataset_loader_fs.h),
// sample split (TrainLoop.h:212-250), the FedNL run on the GPU engine, the
// per-round report and tracking file (TrainLoop.h:753-792, 1655-1862), and
// dopt's shared-library entry doptExecuteLocal (src/mainShared.cpp:15-39).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <string_view>
#include <vector>

#include "fednl_hip.h"

namespace fednl_cli {

// TrainReturnCodes (Args.h:53-64)
enum Rc {
    kOk = 0,
    kMissingArgument = -1,
    kWrongArgument = -2,
    kDatasetNotAvailable = -3,
    kDatasetParse = -4,
    kSplitError = -5,
    kSaveFailed = -6,
    kInternal = -7,
    kUnknownDebugMode = -8
};

// "--name value" / "--flag" lookup with one or two leading dashes ignored
// (dopt/dopt/cmdline/src/CmdLineParser.cpp:129-154)
class CmdLine {
public:
    CmdLine(int argc, char** argv) : argc_(argc), argv_(argv) {}

    bool flag(std::string_view name) const { return find(name) >= 0; }

    bool str(std::string_view name, std::string& out) const
    {
        int i = find(name);
        if (i < 0 || i + 1 >= argc_)
            return false;
        out = argv_[i + 1];
        return true;
    }

    template <class T>
    bool num(std::string_view name, T& out) const
    {
        std::string s;
        if (!str(name, s))
            return false;
        std::istringstream is(s);
        T v;
        if (!(is >> v))
            return false;
        out = v;
        return true;
    }

private:
    static std::string_view strip(std::string_view s)
    {
        if (!s.empty() && s[0] == '-') s.remove_prefix(1);
        if (!s.empty() && s[0] == '-') s.remove_prefix(1);
        return s;
    }
    int find(std::string_view name) const
    {
        for (int i = 1; i < argc_; ++i) {
            if (argv_[i][0] != '-')
                continue;
            if (strip(argv_[i]) == strip(name))
                return i;
        }
        return -1;
    }
    int argc_;
    char** argv_;
};

struct Args {
    // dataset (Args.h:135-170)
    std::string path;
    int clients = 1;
    bool reshuffle = false;
    uint32_t reshuffle_seed = 123;
    bool add_intercept = false;
    // tracking (Args.h:176-205)
    bool tracking = false;
    std::string out;
    // problem
    double lambda = 0.0;
    // runtime (Args.h:226-262)
    int workers = 1, rounds = 0, server_grad_workers = 0, server_hessian_workers = 0;
    uint32_t x0_seed = 1234;
    // algorithm (Args.h:268-395)
    std::string algorithm, compressor;
    double k_mult = 1.0;
    double alpha = 0.0;
    bool theoretical_alpha = false, alpha_opt1 = false, alpha_opt2 = false;
    bool line_search = false;
    double c_ls = 0.5, gamma_ls = 0.5;
    uint32_t max_ls = 5;
    bool option_a = false, option_b = false;
    // GD (Args.h:280-284, 366-379)
    double global_lr = 0.0, global_mult = 1.0;
    bool theoretical_global_lr = false;
    // debug (Args.h:74-129)
    bool silent_print = false, meminfo = false, forceseq = false, crc_data = false, print_train = false;
    bool silent = false;
    bool transfer_idx = false;
    // FedNL-PP (bin_fednl_local_pp/include/Args.h:242-264)
    bool pp = false;
    int clients_per_round = -1;
};

static int parse_debug(const CmdLine& cl, Args& a)
{
    std::string s;
    if (!cl.str("debug", s))
        return kOk;
    std::string cur;
    auto take = [&](const std::string& mode) -> int {
        if (mode.empty()) return kOk;
        if (mode == "notracingprint") a.silent_print = true;
        else if (mode == "meminfo") a.meminfo = true;
        else if (mode == "forcesequpdate") a.forceseq = true;
        else if (mode == "crc32fordata") a.crc_data = true;
        else if (mode == "printtrain") a.print_train = true;
        else {
            std::cerr << "Unknown debug mode: " << mode << '\n';
            return kUnknownDebugMode;
        }
        return kOk;
    };
    for (char ch : s) {
        if (ch == ',' || ch == ';') {
            int rc = take(cur);
            if (rc) return rc;
            cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    return take(cur);
}

static bool file_exists(const std::string& p)
{
    std::ifstream f(p, std::ios::binary);
    return (bool)f;
}

static int parse_args(const CmdLine& cl, Args& a)
{
    // a.pp is set by the caller (which executable this is)
    int rc = parse_debug(cl, a);
    if (rc != kOk)
        return rc;
    a.tracking = cl.flag("tracking");
    if (!a.tracking)
        std::cout << "Tracking information [IGNORE]\n";
    if (!cl.str("out", a.out)) {
        std::cerr << "You did not specify argument \"out\" with path to results which optionally has tracking information\n";
        return kMissingArgument;
    }
    cl.num("clients", a.clients);
    a.reshuffle = cl.flag("reshuffle_train");
    cl.num("reshuffle_train_seed", a.reshuffle_seed);
    a.add_intercept = cl.flag("add_intercept");
    if (!cl.str("train_dataset", a.path)) {
        std::cerr << "Please specify argument \"train_dataset\" with path to train dataset\n";
        return kMissingArgument;
    }
    if (!file_exists(a.path)) {
        std::cerr << "Used train dataset: '" << a.path << "'. File exist: [NOT]\n";
        return kDatasetNotAvailable;
    }
    if (!cl.num("lambda", a.lambda)) {
        std::cerr << "Please specify argument \"lambda\" with value of L2 regulirization for logistic regression\n";
        return kMissingArgument;
    }
    if (!cl.str("algorithm", a.algorithm)) {
        std::cerr << "Please specify argument \"algorithm\" with name of algorithm\n";
        return kMissingArgument;
    }
    if (a.algorithm != "gd") {
        if (!cl.num("k_compressor_as_d_mult", a.k_mult)) {
            std::cerr << "Please specify argument \"k_compressor_as_d_mult\" with value of multiplier for K=<mult>D for Rand-K\n";
            return kMissingArgument;
        }
        if (!cl.str("compressor", a.compressor)) {
            std::cerr << "Please specify argument \"compressor\" with value [identical,randk,topk,seqk]\n";
            return kMissingArgument;
        }
        a.theoretical_alpha = cl.flag("theoretical_alpha");
        a.alpha_opt1 = cl.flag("use_theoretical_alpha_option_1");
        a.alpha_opt2 = cl.flag("use_theoretical_alpha_option_2");
        if (!a.theoretical_alpha && !cl.num("alpha_step_size", a.alpha)) {
            std::cerr << "Please specify argument \"alpha_step_size\" with value of step size for FedNL\n";
            return kMissingArgument;
        }
        a.line_search = !a.pp && cl.flag("line_search");  // bin_fednl_local_pp has no line search
        if (a.line_search) {
            if (!cl.num("c_line_search", a.c_ls)) {
                std::cerr << "Please specify argument \"c_line_search\" for line search. Line search parameter c in (0, 0.5). \n";
                return kMissingArgument;
            }
            if (!cl.num("gamma_line_search", a.gamma_ls)) {
                std::cerr << "Please specify argument \"gamma_line_search\" for line search. Line search parameter c in (0, 1). \n";
                return kMissingArgument;
            }
            cl.num("max_line_search_steps", a.max_ls);
        }
    } else {
        // GD (Args.h:366-379)
        a.theoretical_global_lr = cl.flag("theoretical_global_lr");
        if (!cl.num("global_lr", a.global_lr) && !a.theoretical_global_lr) {
            std::cerr << "Please specify argument \"global_lr\" with value of global_lr\n";
            return kMissingArgument;
        }
        cl.num("global_step_size_multiplier", a.global_mult);
    }
    cl.num("server_grad_workers", a.server_grad_workers);
    cl.num("server_hessian_workers", a.server_hessian_workers);
    cl.num("workers", a.workers);
    if (!cl.num("rounds", a.rounds)) {
        std::cerr << "Please specify argument \"rounds\" with number of rounds\n";
        return kMissingArgument;
    }
    cl.num("x0_seed", a.x0_seed);
    if (a.pp)
        cl.num("clients-per-round", a.clients_per_round);
    a.option_a = cl.flag("fednl-option-a");
    a.option_b = cl.flag("fednl-option-b");
    a.silent = cl.flag("silent");
    a.transfer_idx = cl.flag("transfer_indicies_for_randk");
    return kOk;
}

// ---------------------------------------------------------------------------
// LIBSVM text, read as dopt's TextFileDataSetLoader reads it for train()
// (TrainLoop.h:123-170; dataset_loader_fs.h:381-604 statistics pass,
// 606-883 createDataset): lines end at '\n' or '\r', fields are split on ' '
// or '\t', and each field on ':' (empty pieces dropped). A field of one piece
// is a label, of two pieces a feature index:value pair.
//  - Statistics pass: labels parse as int32 and indices as int32, both with
//    dopt's strict decimal parser. On the first example every label resets
//    the label range (and the class count to 1) and every index resets the
//    index range; later, a label equal to neither end of the range is a new
//    class (and widens the range if outside it), and an index outside the
//    range widens it. Exactly two classes are required.
//  - Create pass: d = max index - min index + 1 (+ 1 intercept, the last
//    column); labels parse as double and map to -1 below the mean of the
//    label range's ends, +1 otherwise; values parse as double with dopt's
//    decimal parser (integer part + fraction part * 10^-length, scaled by a
//    power of ten for an exponent), which can differ from strtod in the last
//    bit.
// ---------------------------------------------------------------------------
struct Dataset {
    int N = 0, d = 0;
    std::vector<double> X;  // N x d, row-major
    std::vector<double> y;  // N
};

// StringUtils.h:171-268 (fromStringSignedInteger): optional sign, then
// decimal digits only; int32 arithmetic (wraps)
static bool parse_i32(std::string_view s, int32_t& out)
{
    if (s.empty())
        return false;
    size_t p = 0;
    int sign = 1;
    if (s[0] == '+' || s[0] == '-') {
        sign = s[0] == '-' ? -1 : 1;
        p = 1;
    }
    if (p >= s.size() || s[p] < '0' || s[p] > '9')
        return false;
    uint32_t v = (uint32_t)(s[p] - '0');
    if (sign < 0) v = 0u - v;
    for (++p; p < s.size(); ++p) {
        if (s[p] < '0' || s[p] > '9')
            return false;
        const uint32_t dgt = (uint32_t)(s[p] - '0');
        v = sign > 0 ? v * 10u + dgt : v * 10u - dgt;
    }
    out = (int32_t)v;
    return true;
}

// SimpleMathRoutines.h:179-195: a^n by halving the exponent
static double power_natural(double a, uint32_t n)
{
    if (n == 0)
        return 1.0;
    const double t = power_natural(a, n / 2);
    return (n % 2 == 0) ? t * t : t * t * a;
}

// StringUtils.h:270-463 (fromStringFloat) for double
static bool parse_dopt_double(std::string_view s, double& out)
{
    if (s.empty())
        return false;
    size_t p = 0;
    int sign = 1;
    if (s[0] == '+') {
        p = 1;
    } else if (s[0] == '-') {
        sign = -1;
        p = 1;
    }
    uint32_t ip = 0, fp = 0, flen = 0, ep = 0;
    int esign = 1;
    for (; p < s.size(); ++p) {
        const char c = s[p];
        if (c == '.' || c == 'e' || c == 'E')
            break;
        if (c < '0' || c > '9')
            return false;
        ip = ip * 10u + (uint32_t)(c - '0');
    }
    if (p < s.size() && s[p] == '.') {
        for (++p; p < s.size(); ++p) {
            const char c = s[p];
            if (c == 'e' || c == 'E')
                break;
            if (c < '0' || c > '9')
                return false;
            fp = fp * 10u + (uint32_t)(c - '0');
            ++flen;
        }
    }
    if (p < s.size() && (s[p] == 'e' || s[p] == 'E')) {
        ++p;
        if (p < s.size() && s[p] == '+') {
            ++p;
        } else if (p < s.size() && s[p] == '-') {
            esign = -1;
            ++p;
        }
        for (; p < s.size(); ++p) {
            const char c = s[p];
            if (c < '0' || c > '9')
                return false;
            ep = ep * 10u + (uint32_t)(c - '0');
        }
    }
    const double I = (double)(int32_t)ip, F = (double)(int32_t)fp;
    if (ep == 0) {
        if (flen == 0) {
            out = sign > 0 ? I : -I;
        } else {
            // the reference's table holds 1 / 1e0 .. 1 / 1e16 as constants
            static const double kDen[] = {1.0 / 1e0,  1.0 / 1e1,  1.0 / 1e2,  1.0 / 1e3,  1.0 / 1e4,  1.0 / 1e5,
                                          1.0 / 1e6,  1.0 / 1e7,  1.0 / 1e8,  1.0 / 1e9,  1.0 / 1e10, 1.0 / 1e11,
                                          1.0 / 1e12, 1.0 / 1e13, 1.0 / 1e14, 1.0 / 1e15, 1.0 / 1e16};
            const double den = flen < 17 ? kDen[flen] : 1.0 / power_natural(10.0, flen);
            const double v = I + F * den;
            out = sign > 0 ? v : -v;
        }
    } else {
        const double den = 1.0 / power_natural(10.0, flen);
        const double e = power_natural(10.0, ep);
        const double m = I + F * den;
        if (esign > 0)
            out = sign > 0 ? m * e : -m * e;
        else
            out = sign > 0 ? m / e : -m / e;
    }
    return true;
}

// splitToSubstrings without empty pieces (StringUtils.h:667-716)
template <class Pred>
static void split_pieces(std::string_view s, Pred is_sep, std::vector<std::string_view>& out)
{
    out.clear();
    size_t a = 0;
    for (size_t i = 0; i <= s.size(); ++i) {
        if (i == s.size() || is_sep(s[i])) {
            if (i > a)
                out.emplace_back(s.substr(a, i - a));
            a = i + 1;
        }
    }
}

// the text's non-empty lines, in order
static void text_lines(const std::string& text, std::vector<std::string_view>& lines)
{
    split_pieces(std::string_view(text), [](char c) { return c == '\n' || c == '\r'; }, lines);
}

static int load_libsvm(const std::string& path, bool add_intercept, Dataset& ds, bool print_errors = true)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return kDatasetParse;
    const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<std::string_view> lines, fields, kv;
    text_lines(text, lines);
    auto field_sep = [](char c) { return c == ' ' || c == '\t'; };
    auto kv_sep = [](char c) { return c == ':'; };
    // statistics pass
    int32_t lmin = 0, lmax = 0, imin = 0, imax = 0;
    int unique_labels = 0;
    size_t examples = 0;
    for (std::string_view line : lines) {
        split_pieces(line, field_sep, fields);
        if (fields.empty())
            continue;
        for (std::string_view fld : fields) {
            split_pieces(fld, kv_sep, kv);
            if (kv.size() == 1) {
                int32_t lv = 0;
                if (!parse_i32(kv[0], lv)) {
                    if (print_errors)
                        std::cerr << "Error during conversion of lable: " << kv[0] << ". Parsing Line: " << examples + 1
                                  << '\n';
                } else if (examples == 0) {
                    lmin = lmax = lv;
                    unique_labels = 1;
                } else if (lv != lmin && lv != lmax) {
                    ++unique_labels;
                    if (lv < lmin)
                        lmin = lv;
                    else if (lv > lmax)
                        lmax = lv;
                }
            } else if (kv.size() == 2) {
                int32_t idx = 0;
                if (!parse_i32(kv[0], idx)) {
                    if (print_errors)
                        std::cerr << "Error during parsing. Line: " << examples + 1 << ". Can not parse " << kv[0]
                                  << " for purpose of converting to unsigned integer.\n";
                    return kDatasetParse;
                }
                if (examples == 0) {
                    imin = imax = idx;
                } else if (idx < imin) {
                    imin = idx;
                } else if (idx > imax) {
                    imax = idx;
                }
            } else if (print_errors) {
                std::cerr << "Warning during parsing. Line: " << examples + 1 << ". Can not parse " << fld
                          << " for purpose of obtaing feature_index:value of label.\n";
            }
        }
        ++examples;
    }
    if (examples == 0)
        return kDatasetParse;
    if (unique_labels != 2) {
        std::cout << "Specified dataset labels for " << unique_labels
                  << " classes. And you need to have a dataset with 2 classes.\n";
        return kDatasetParse;
    }
    // create pass
    const int dfeat = (int)((int64_t)imax - (int64_t)imin + 1);
    ds.N = (int)examples;
    ds.d = dfeat + (add_intercept ? 1 : 0);
    ds.X.assign((size_t)ds.N * ds.d, 0.0);
    ds.y.assign(ds.N, 0.0);
    const double avg = ((double)lmin + (double)lmax) / 2.0;
    size_t s = 0;
    for (std::string_view line : lines) {
        split_pieces(line, field_sep, fields);
        if (fields.empty())
            continue;
        double* row = ds.X.data() + s * ds.d;
        for (std::string_view fld : fields) {
            split_pieces(fld, kv_sep, kv);
            if (kv.size() == 1) {
                double lv = 0.0;
                if (parse_dopt_double(kv[0], lv))
                    ds.y[s] = lv;
                else if (print_errors)
                    std::cerr << "Error during parsing. Line: " << s + 1 << ". Can not parse " << kv[0]
                              << " for purpose of converting to response variable.\n";
                ds.y[s] = (ds.y[s] < avg) ? -1.0 : 1.0;
            } else if (kv.size() == 2) {
                int32_t idx = 0;
                double v = 0.0;
                const bool ok_i = parse_i32(kv[0], idx);
                const bool ok_v = parse_dopt_double(kv[1], v);
                if (!ok_i || !ok_v) {
                    if (print_errors)
                        std::cerr << "Error during parsing. Line: " << s + 1 << ". Can not parse: " << fld << '\n';
                    continue;  // the reference stores an unparsed index or value uninitialised
                }
                const int64_t col = (int64_t)idx - imin;
                if (col >= 0 && col < dfeat)  // outside: the reference writes out of its matrix
                    row[col] = v;
            } else if (print_errors) {
                std::cerr << "Error during parsing. Line: " << s + 1
                          << ". More then one feature_index:value pair has been found.\n";
            }
        }
        if (add_intercept)
            row[ds.d - 1] = 1.0;
        ++s;
    }
    return kOk;
}

// Fisher-Yates of Shuffle.h:101-158 driven by RandomGenIntegerLinear
// (RandomGenIntegerLinear.h:49-58): u = x / (2^32 - 1), offset rounded as
// SimpleMathRoutines.h:407-414.
static void shuffle_indices(std::vector<int32_t>& v, uint32_t seed)
{
    uint32_t x = seed;
    size_t rem = v.size() - 1;
    for (size_t i = 0; i < v.size(); ++i, --rem) {
        x = 214013u * x + 2531011u;
        const double u = (double)x * (1.0 / 4294967295.0);
        const double p = u * (double)rem;
        size_t off = (size_t)p;
        if (p - (double)off > 0.5) off += 1;
        if (off != 0) std::swap(v[i], v[i + off]);
    }
}

// CRC-32 of dopt::crc32 (dopt/dopt/system/src/digest/Crc.cpp:74-90): reflected
// polynomial 0xEDB88320, seed 0xFFFFFFFF, no final inversion.
static uint32_t crc32_bytes(const void* buf, size_t size, uint32_t crc = 0xFFFFFFFFu)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int b = 0; b < 8; ++b) c = (c & 1) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
            table[i] = c;
        }
        init = true;
    }
    const unsigned char* p = (const unsigned char*)buf;
    for (size_t i = 0; i < size; ++i) crc = (crc >> 8) ^ table[(crc ^ p[i]) & 0xFFu];
    return crc;
}

static std::string base_name(const std::string& p)
{
    const size_t s = p.find_last_of("/\\");
    return s == std::string::npos ? p : p.substr(s + 1);
}

struct Bytes {
    std::vector<char> b;
    void str(const std::string& s, bool zero) { b.insert(b.end(), s.begin(), s.end()); if (zero) b.push_back(0); }
    void i32(int32_t v) { const char* p = (const char*)&v; b.insert(b.end(), p, p + 4); }
    void f64(double v) { const char* p = (const char*)&v; b.insert(b.end(), p, p + 8); }
};

// dataset_loader_fs.h:237-313 (checkSplitConsistency) on this split: the
// clients' sample groups are disjoint, every client row is the training row it
// was copied from (features and label), and the clients' samples and the
// skipped ones together cover every sample.
static bool check_split(const Dataset& ds, const std::vector<int32_t>& order, int n, int mpc,
                        const std::vector<double>& feats, const std::vector<double>& labs)
{
    bool ok = true;
    const int d = ds.d;
    std::vector<char> seen(ds.N, 0);
    for (int c = 0; c < n; ++c) {
        bool repeat = false;
        for (int s = 0; s < mpc; ++s) {
            const int g = order[(size_t)c * mpc + s];
            if (g < 0 || g >= ds.N || seen[g])
                repeat = true;
            else
                seen[g] = 1;
        }
        if (repeat) {
            std::cerr << "Error in splitting. Some indicies are repeating for group #" << c << '\n';
            ok = false;
        }
    }
    for (int c = 0; c < n; ++c) {
        bool bad = false;
        for (int s = 0; s < mpc; ++s) {
            const int g = order[(size_t)c * mpc + s];
            if (g < 0 || g >= ds.N) {
                bad = true;
                continue;
            }
            const size_t r = (size_t)c * mpc + s;
            if (labs[r] != ds.y[g] || memcmp(&feats[r * d], &ds.X[(size_t)g * d], sizeof(double) * d) != 0)
                bad = true;
        }
        if (bad) {
            std::cerr << "Error in splitting. Something has not been copied in a correct way for group #" << c << '\n';
            ok = false;
        }
    }
    for (size_t q = (size_t)n * mpc; q < order.size(); ++q)
        if (order[q] >= 0 && order[q] < ds.N)
            seen[order[q]] = 1;
    for (int g = 0; g < ds.N; ++g)
        if (!seen[g]) {
            std::cerr << "Error in splitting. The record " << g << " has been lost\n";
            ok = false;
        }
    return ok;
}

// logistic_regression.h:707-740 (computeLSmoothness with tolerance 1e-3 mu,
// TrainLoop.h:314): power iteration on A^T A / (4 M) over the clients' M
// samples from a RandomGenRealLinear start (seed 1234, [0, 1],
// RandomGenRealLinear.h:73-77; VectorND_Raw.h:835-840), normalised by the
// infinity norm; L = the last norm + lambda. A^T A / (4 M) is the average of
// the clients' Hessians at x = 0 (sigma(0)(1 - sigma(0)) = 1/4) less lambda I,
// formed by the engine.
static bool l_smoothness(void* eng, int n, int d, double lambda, double& L, size_t& iters)
{
    const size_t S = (size_t)d * (d + 1) / 2;
    std::vector<double> hc((size_t)n * S), x0(d, 0.0);
    if (fednl_hessians_at(eng, x0.data(), hc.data()) != FEDNL_OK)
        return false;
    std::vector<double> B((size_t)d * d);
    for (int j = 0; j < d; ++j)
        for (int i = 0; i <= j; ++i) {
            const size_t p = (size_t)j * (j + 1) / 2 + i;
            double acc = 0.0;
            for (int c = 0; c < n; ++c) acc += hc[(size_t)c * S + p];
            double v = acc / n;
            if (i == j) v -= lambda;
            B[(size_t)i * d + j] = B[(size_t)j * d + i] = v;
        }
    uint32_t xs = 1234u;
    std::vector<double> v(d), vp(d);
    for (int i = 0; i < d; ++i) {
        xs = (uint32_t)((214013ull * xs + 2531011ull) % 4294967296ull) % 4294967295u;
        v[i] = 0.0 + ((double)xs * (1.0 / 4294967295.0)) * (1.0 - 0.0);
    }
    auto linf = [](const std::vector<double>& u) {
        double r = fabs(u[0]);
        for (size_t i = 1; i < u.size(); ++i) r = std::max(r, fabs(u[i]));
        return r;
    };
    vp = v;
    const double v0 = linf(vp);
    for (auto& t : vp) t /= v0;
    const double eps = 0.001 * lambda;
    double maxeig = 0.0;
    for (iters = 0;; ++iters) {
        for (int i = 0; i < d; ++i) {
            double acc = 0.0;
            for (int j = 0; j < d; ++j) acc += B[(size_t)i * d + j] * vp[j];
            v[i] = acc;
        }
        maxeig = linf(v);
        for (auto& t : v) t /= maxeig;
        double diff = 0.0;
        for (int i = 0; i < d; ++i) diff = std::max(diff, fabs(v[i] - vp[i]));
        if (diff < eps || iters > 100000)
            break;
        std::swap(v, vp);
    }
    L = maxeig + lambda;
    return true;
}

static int train(const CmdLine& cl, ResultCallbackLocal cb, ResultCallbackLocalPP cb_pp, bool pp)
{
    Args a;
    a.pp = pp;
    int rc = parse_args(cl, a);
    if (rc != kOk)
        return rc;
    const bool print = !a.silent_print;

    Dataset ds;
    rc = load_libsvm(a.path, a.add_intercept, ds);
    if (rc != kOk) {
        if (print) std::cerr << "Train dataset: '" << a.path << "'. [Error during parsing]\n";
        return kDatasetParse;
    }
    if (print) {
        std::cout << "Train dataset: '" << a.path << "' was instantiated [OK].\n";
        std::cout << "About data set: \"train dataset\"\n";
        std::cout << "  Total number of samples in data set: " << ds.N << '\n';
        std::cout << "  Total number of attributes in data set: " << ds.d << '\n';
        std::cout << "  Design matrix includes intercept term: " << (a.add_intercept ? "[YES]" : "[NO]") << '\n';
        if (a.crc_data) {
            // the reference hashes its d x N column-major design (padded to LDA) and labels
            const int lda = (ds.d + 3) / 4 * 4;
            std::vector<double> col((size_t)lda * ds.N, 0.0);
            for (int s = 0; s < ds.N; ++s)
                for (int i = 0; i < ds.d; ++i) col[(size_t)s * lda + i] = ds.X[(size_t)s * ds.d + i];
            std::cout << "  Dataset CRC-32 samples: " << crc32_bytes(col.data(), col.size() * 8) << '\n';
            std::cout << "  Dataset CRC-32 labels: " << crc32_bytes(ds.y.data(), ds.y.size() * 8) << '\n';
        }
    }
    const int n = a.clients;
    if (n <= 0 || n > ds.N) {
        std::cerr << "Number of clients must be in [1, number of samples]\n";
        return kWrongArgument;
    }
    if (a.pp && (a.clients_per_round < 1 || a.clients_per_round > n)) {
        // bin_fednl_local_pp asserts 0 < clients-per-round <= clients (TrainLoop.h:282)
        std::cerr << "Please specify --clients-per-round in [1, " << n << "]\n";
        return kWrongArgument;
    }
    const int mpc = ds.N / n;  // TrainLoop.h:212-250: residual samples are dropped
    std::vector<int32_t> order(ds.N);
    for (int s = 0; s < ds.N; ++s) order[s] = s;
    if (a.reshuffle)
        shuffle_indices(order, a.reshuffle_seed);
    const int d = ds.d;
    // client-major design for the engine: client c takes order[c m .. (c+1) m)
    std::vector<double> feats((size_t)n * mpc * d), labs((size_t)n * mpc);
    for (int c = 0; c < n; ++c)
        for (int s = 0; s < mpc; ++s) {
            const int src = order[(size_t)c * mpc + s];
            memcpy(&feats[((size_t)c * mpc + s) * d], &ds.X[(size_t)src * d], sizeof(double) * d);
            labs[(size_t)c * mpc + s] = ds.y[src];
        }
    if (print) {
        std::cout << "  Train samples has been reshuffled: " << (a.reshuffle ? "[YES]" : "[NO]") << '\n';
        std::cout << "  Train samples reshuffled seed: " << a.reshuffle_seed << '\n';
        std::cout << "  Number of clients: " << n << '\n';
        std::cout << "  Number of samples per client: " << mpc << '\n';
        std::cout << "  Number of skipped samples during samples distribution: " << (ds.N - mpc * n) << '\n';
        std::cout << "  Train data distribution across clients has been completed. [OK]\n";
        std::cout << "=========================================================================\n";
    }
    if (cl.flag("check_split")) {
        if (!check_split(ds, order, n, mpc, feats, labs)) {
            if (print) std::cerr << "  Checking data split [FAILED]\n";
            return kSplitError;
        }
        if (print) std::cout << "  Checking data split [OK]\n";
    } else if (print) {
        std::cout << "  Checking data split [SKIP] \n";
    }
    if (print)
        std::cout << "=========================================================================\n";

    // algorithm selector: gd || fednl1 (TrainLoop.h:297-330)
    const bool gd = a.algorithm == "gd";
    const bool fednl = a.algorithm == "fednl1";
    if (fednl && !a.option_a && !a.option_b) {
        if (print) std::cerr << "For FedNL you should specify fednl-option-a or fednl-option-b\n";
        return kMissingArgument;
    }
    // the compressor (TrainLoop.h:367-400; GD has none)
    int comp = FEDNL_RANDSEQK;
    if (!gd) {
        if (a.compressor == "identical") comp = FEDNL_IDENTICAL;
        else if (a.compressor == "natural") comp = FEDNL_NATURAL;
        else if (a.compressor == "randk") comp = FEDNL_RANDK;
        else if (a.compressor == "topk") comp = FEDNL_TOPK;
        else if (a.compressor == "toplek") comp = FEDNL_TOPLEK;
        else if (a.compressor == "seqk") comp = FEDNL_RANDSEQK;
        else {
            if (print) std::cerr << "You have specified unknown type of compressor: " << a.compressor << '\n';
            return kWrongArgument;
        }
    }
    if (!gd && !fednl) {
        // TrainLoop.h:653-658
        if (print) std::cerr << " Please specify valid algorithm name.\n";
        return kInternal;
    }
    if (print) {
        std::cout << "Main Loop preparation for: " << a.algorithm << '\n';
        std::cout << "=========================================================================\n";
    }
    const size_t S = (size_t)d * (d + 1) / 2;
    const bool whole = !gd && (comp == FEDNL_IDENTICAL || comp == FEDNL_NATURAL);  // TrainLoop.h:354-377
    const int k = whole ? (int)S : (int)(a.k_mult * d);
    if (!whole && (size_t)k > S) {
        if (print)
            std::cerr << " Multiplier for D is specifed as " << a.k_mult << ". For this problem with D=" << d
                      << " maximum allowable muliplier is " << double(S) / double(d) << "!\n";
        return kWrongArgument;
    }
    int alpha_mode = 0;
    if (a.theoretical_alpha && !gd) {
        if (comp == FEDNL_TOPK || comp == FEDNL_TOPLEK) {
            if (a.alpha_opt1) alpha_mode = 1;
            else if (a.alpha_opt2) alpha_mode = 2;
            else {
                if (print)
                    std::cerr << "Please specify --use_theoretical_alpha_option_2 or --use_theoretical_alpha_option_1 in command line\n";
                return kMissingArgument;
            }
        } else {
            alpha_mode = 2;
        }
    }

    fednl_opts o;
    fednl_default_opts(&o);
    o.lambda = a.lambda;
    o.compressor = comp;
    o.k_mult = a.k_mult;
    o.alpha_mode = alpha_mode;
    o.alpha = a.alpha;
    o.option_b = a.option_b ? 1 : 0;
    o.workers = a.workers;
    o.tracking = a.tracking ? 1 : 0;
    o.line_search = a.line_search ? 1 : 0;
    o.c_line_search = a.c_ls;
    o.gamma_line_search = a.gamma_ls;
    o.max_line_search_steps = (int)a.max_ls;
    o.clients_per_round = a.pp ? a.clients_per_round : 0;
    o.algorithm = gd ? FEDNL_ALG_GD : FEDNL_ALG_FEDNL;
    // TrainLoop.h:358, 521 (the theoretical step is set once L is known)
    double used_global_step = a.global_lr * a.global_mult;
    o.global_lr = used_global_step;
    void* eng = nullptr;
    rc = fednl_create(&o, n, mpc, d, feats.data(), labs.data(), &eng);
    if (rc != FEDNL_OK) {
        std::cerr << "engine: " << fednl_last_error() << '\n';
        return kInternal;
    }
    // GD has no alpha: the reported one is --alpha_step_size's default
    const double alpha = gd ? a.alpha : fednl_alpha(eng);
    // TrainLoop.h:306-316: L(f) when asked for, or for GD's theoretical
    // global step size (a flag that only GD parses)
    const bool compute_L = cl.flag("compute-L-smooth") || (gd && a.theoretical_global_lr);
    double L_f = -1.0;
    if (compute_L) {
        size_t iters = 0;
        if (!l_smoothness(eng, n, d, a.lambda, L_f, iters)) {
            std::cerr << "engine: " << fednl_last_error() << '\n';
            fednl_destroy(eng);
            return kInternal;
        }
        if (print)
            std::cout << "Compute L constants for objective (power iterations: " << iters << ")\n";
    }
    if (gd && a.theoretical_global_lr) {
        // TrainLoop.h:403-407, 521: 2 / (mu + L) times the multiplier
        used_global_step = 2.0 / (a.lambda + L_f);
        used_global_step *= a.global_mult;
        if (fednl_set_global_lr(eng, used_global_step) != FEDNL_OK) {
            std::cerr << "engine: " << fednl_last_error() << '\n';
            fednl_destroy(eng);
            return kInternal;
        }
    }
    if (print) {
        std::cout << "=========================================================================\n";
        std::cout << "Main loop has been started: r:" << a.rounds << ",clients:" << n << ",d:" << d << ",ni:" << mpc << '\n';
    }
    rc = fednl_enqueue(eng, a.rounds);
    if (rc == FEDNL_OK)
        rc = fednl_sync(eng);
    if (rc != FEDNL_OK) {
        std::cerr << "engine: " << fednl_last_error() << '\n';
        fednl_destroy(eng);
        return kInternal;
    }
    int R = 0;
    fednl_trace(eng, nullptr, nullptr, nullptr, nullptr, nullptr, &R);
    std::vector<double> gn(R), fv(R), xn(R), sec(R);
    std::vector<int> ok(R);
    fednl_trace(eng, gn.data(), fv.data(), xn.data(), sec.data(), ok.data(), &R);
    std::vector<int> cnt(n);
    std::vector<double> x(d), glast(d);
    fednl_get_x(eng, x.data());
    fednl_last_round(eng, glast.data(), nullptr, cnt.data(), nullptr, nullptr, nullptr, nullptr);
    std::vector<int> ls_trials(R, 0), parts(R, n);
    std::vector<double> items(R, 0.0);
    fednl_trace_line_search(eng, ls_trials.data(), &R);
    fednl_trace_clients(eng, parts.data(), items.data(), &R);

    // bytes received from the clients per round (TrainLoop.h:1104,1131-1132,1287,1533,1593):
    // every participating client (all n, or tau for FedNL-PP) sends its
    // gradient (8d), the Hessian items and, with option b, Lk (8); the f_i
    // message of the round is counted with the size of the Lk message (8 with
    // option b, none with option a); each line-search trial adds one 8-byte
    // f_i per client. Items: k per client, TopLEK its counts, Natural three
    // bytes per pair of values with one trailing pair
    // (LocalWorkerForFedNL_Natural.h:80-120)
    const bool send_fi = a.tracking || a.line_search;  // TrainLoop.h:620
    auto round_bytes = [&](int r, double& scal, double& idxb) {
        if (gd) {
            // every client's gradient (TrainLoop.h:1022)
            scal = (double)n * 8.0 * d;
            idxb = 0.0;
            return;
        }
        const double np = (double)parts[r];
        const double item_bytes = (comp == FEDNL_NATURAL) ? np * 3.0 * (double)(S / 2 + 1) : 8.0 * items[r];
        scal = np * (8.0 * d + (a.option_b ? 8.0 : 0.0) + ((send_fi && a.option_b) ? 8.0 : 0.0) +
                     8.0 * ls_trials[r]) + item_bytes;
        idxb = 0.0;
        if (comp == FEDNL_TOPK) idxb = np * 4.0 * k;
        else if (comp == FEDNL_TOPLEK) idxb = 4.0 * np + 4.0 * items[r];
        else if ((comp == FEDNL_RANDK || comp == FEDNL_RANDSEQK) && a.transfer_idx) idxb = np * 4.0 * k;
    };
    Bytes tr;
    if (a.tracking) {
        tr.str(base_name(a.path), true);
        std::string alg = a.pp ? "FedNL-PP" : "FedNL";
        if (a.line_search) alg += "-LS";
        if (a.option_a) alg += " (a)";
        if (a.option_b) alg += " (b)";
        if (gd) alg = "GD";  // TrainLoop.h:767-770
        tr.str(alg, true);
        tr.str(a.compressor, true);
        tr.i32(mpc);
        tr.i32(ds.N);
        tr.i32(a.rounds);
        tr.i32(n);
        tr.i32(d);
        tr.i32(k);
        tr.f64(a.k_mult);
        tr.f64(used_global_step);
        tr.f64(alpha);
        tr.f64(a.lambda);
    }
    double acc_s = 0.0, acc_i = 0.0;
    for (int r = 0; r < R; ++r) {
        // the master reports the fallback solve before the round line
        if (!ok[r] && print)
            std::cout << "SST FACTORIZATION FAILED\n";
        if (!a.silent && print) {
            const double t = sec[r] > 0 ? sec[r] : 1e-9;
            std::cout << "  Round #" << r << ": norm of full gradient: " << gn[r];
            if (a.pp) {
                // bin_fednl_local_pp/include/TrainLoop.h:1688-1696
                std::cout << ": f(x): " << fv[r] << ": avg. time per round: " << 1e3 * t / (r + 1)
                          << " milliseconds" << '\n';
            } else {
                if (send_fi) std::cout << ": f(x): " << fv[r];
                else std::cout << ": f(x): [N/A]";
                std::cout << ": avg. time per round: " << 1e3 * t / (r + 1) << " msec"
                          << ": rounds/second: " << (r + 1) / t << ": line-search num: " << ls_trials[r] << '\n';
            }
        }
        if (a.tracking) {
            tr.i32(r);
            tr.f64(gn[r]);
            tr.f64(fv[r]);
            tr.f64(xn[r]);
            tr.f64(acc_s);
            tr.f64(acc_i);
            tr.f64(sec[r]);
        }
        double bs, bi;
        round_bytes(r, bs, bi);
        acc_s += bs;
        acc_i += bi;
    }
    const double elapsed_ms = R > 0 ? sec[R - 1] * 1e3 : 0.0;
    double gg = 0.0, xx = 0.0;
    for (int i = 0; i < d; ++i) {
        gg += glast[i] * glast[i];
        xx += x[i] * x[i];
    }
    const double mu = a.lambda;  // computeMuStrongConvexity (logistic_regression.h:702-705)
    const double f_gap = gg / (2.0 * mu);   // ConvexOptimizationRoutines.h:12-27
    const double x_gap = 2.0 * gg / mu;
    if (print) {
        // TrainLoop.h:1700-1712
        std::cout << "=========================================================================\n";
        std::cout << "Main loop has been finished: r:" << a.rounds << ",clients:" << n << ",d:" << d << ",ni:" << mpc
                  << '\n';
        std::cout << "=========================================================================\n";
        std::cout << "Statistics\n\n";
        std::cout << "  Average time per round: " << (R ? elapsed_ms / R : 0.0) << " milliseconds\n";
        std::cout << "  Total number of rounds: " << a.rounds << '\n';
        std::cout << "  Total number of clients: " << n << '\n';
        if (a.pp)
            std::cout << "  Number of clients per round selected u.a.r: " << a.clients_per_round << '\n';
        if (!a.pp) {
            // TrainLoop.h:1716-1720
            std::cout << '\n';
            std::cout << "  Total number of workers for clients: " << a.workers << '\n';
            std::cout << "  Total number of workers for master [for gradient updates]: " << a.server_grad_workers
                      << '\n';
            std::cout << "  Total number of workers for master [for hessian updates]: " << a.server_hessian_workers
                      << '\n';
            std::cout << '\n';
        }
        std::cout << "  Algorithm: " << a.algorithm << '\n';
        std::cout << "  Data set: " << a.path << '\n';
        std::cout << "  d: " << d << '\n';
        if (!a.pp) {
            // TrainLoop.h:1724-1729
            std::cout << '\n';
            std::cout << "  Global step size (with multiplier): " << used_global_step << '\n';
            std::cout << "  Global step size multiplier: " << a.global_mult << '\n';
            std::cout << "  Is global step size theoretical: " << (a.theoretical_global_lr ? "[YES]" : "[NO]") << '\n';
        }
        std::cout << "  Alpha step size: " << alpha << '\n';
        if (!a.pp)
            std::cout << "  Is alpha step size theoretical: " << (a.theoretical_alpha && !gd ? "[YES]" : "[NO]") << '\n';
        std::cout << "  Lambda for regularization: " << a.lambda << '\n';
        std::cout << "  Received MBytes from the clients: " << (acc_s + acc_i) / (1024.0 * 1024.0) << '\n';
        std::cout << "  Clients compressor: " << a.compressor << '\n';
        std::cout << "  Clients compressor K: " << k << " / Hessian shape is [d,d] where d: " << d
                  << " / Maximum K ( {d(d+1)}/2 ): " << S << '\n';
        if (!a.pp) {
            std::cout << '\n';
            std::cout << "  Line Search: " << (a.line_search ? "[YES]" : "[NO]") << '\n';
            std::cout << "  Line Search parameters: " << "c: " << a.c_ls << ", gamma: " << a.gamma_ls << '\n';
        }
        std::cout << "=========================================================================\n";
        // TrainLoop.h:1741-1764
        std::cout << '\n';
        std::cout << "Debug information\n\n";
        std::cout << "  Transfer indicies for RandK: " << (a.transfer_idx ? "[YES]" : "[NO]") << '\n';
        std::cout << "  mu(f): " << mu << '\n';
        if (compute_L) {
            std::cout << "  L(f): " << L_f << '\n';
            std::cout << "  cond(f) = (L(f)/mu(f)): " << L_f / mu << '\n';
        } else {
            std::cout << "  L(f): " << "unavailable" << '\n';
            std::cout << "  cond(f) = (L(f)/mu(f)): " << "unavailable" << '\n';
        }
        std::cout << "  Maximum allowable k multiplier for RandK and TopK: " << double(((d + 1) * d) / 2) / d << '\n';
        std::cout << "  Maximum sendable components for RandK and TopK from one client: " << double(((d + 1) * d) / 2)
                  << '\n';
        std::cout << '\n';
        std::cout << "  Print information about CRC32 for datasets: " << (a.crc_data ? 1 : 0) << '\n';
        std::cout << "  Print information about used memory: " << (a.meminfo ? 1 : 0) << '\n';
        std::cout << "  Server obtainining updates from clients in order: " << (a.forceseq ? 1 : 0) << '\n';
        std::cout << '\n';
        std::cout << "=========================================================================\n";
        std::cout << "Solution Properties\n\n";
        std::cout << "  Last iterate CRC-32: " << crc32_bytes(x.data(), sizeof(double) * d) << '\n';
        std::cout << "  Gradient in last iterate CRC-32: " << crc32_bytes(glast.data(), sizeof(double) * d) << '\n';
        std::cout << "  Norm of last iterate: " << sqrt(xx) << '\n';
        if (a.pp)
            std::cout << "  True of norm of full gradient in last iterate: " << sqrt(gg) << '\n';
        else
            std::cout << "  Norm of full gradient in last iterate: " << sqrt(gg) << '\n';
        std::cout << "  Norm square of full gradient in last iterate: " << gg << '\n';
        std::cout << "  Suboptimality gap of Function, f(x) - f*: " << f_gap << '\n';
        std::cout << "  Suboptimality distance to Optimal Solution |x - x*|: " << x_gap << '\n';
        std::cout << "=========================================================================\n";
        std::cout << "  Training has been finished successfully [OK]\n";
    }
    fednl_destroy(eng);

    // {run_id},{algorithm},{dataset},{kmultiplier} (TrainLoop.h:1813-1831)
    std::string outp = a.out;
    {
        srand((unsigned)time(nullptr));
        const std::string rid = std::to_string((int)((double)rand() / RAND_MAX * 10000));
        const std::string keys[4] = {"{run_id}", "{algorithm}", "{dataset}", "{kmultiplier}"};
        const std::string vals[4] = {rid, a.algorithm, base_name(a.path), std::to_string(a.k_mult)};
        for (int q = 0; q < 4; ++q) {
            const size_t p = outp.find(keys[q]);
            if (p != std::string::npos) outp.replace(p, keys[q].size(), vals[q]);
        }
    }
    if (!outp.empty()) {
        std::ofstream f(outp, std::ios::binary);
        bool saved = (bool)f;
        if (saved && !tr.b.empty()) {
            f.write(tr.b.data(), (std::streamsize)tr.b.size());
            saved = (bool)f;
        }
        if (print)
            std::cout << "  Tracking results has been saved to: " << outp << (saved ? " [OK]" : " [FAILED]") << '\n';
    }
    if (cb)
        cb(d, sqrt(gg), f_gap, x_gap);
    if (cb_pp)
        cb_pp(d);
    return kOk;
}

// "train " + command, split on single spaces (src/mainShared.cpp:15-39 of
// both shared libraries); argv keeps pointers into `parts`
static std::vector<char*> split_command(const char* command, const char* prog, std::vector<std::string>& parts)
{
    std::string full = std::string("train ") + (command ? command : "");
    std::string cur;
    for (char ch : full) {
        if (ch == ' ') {
            parts.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    parts.push_back(cur);
    std::vector<char*> argv;
    argv.push_back(const_cast<char*>(prog));
    for (auto& p : parts) {
        if (p == "train") p = "--train";
        argv.push_back(p.data());
    }
    return argv;
}

static int cli_main(int argc, char** argv, bool pp)
{
    CmdLine cl(argc, argv);
    if (cl.flag("version")) {
        std::cout << "fednl-mi355x engine (FedNL round engine for gfx950)\n";
        if (!cl.flag("train"))
            return 0;
    }
    if (!cl.flag("train")) {
        std::cout << "Please specify one of the action [train]\n";
        return -1;
    }
    return train(cl, nullptr, nullptr, pp);
}

}  // namespace fednl_cli

extern "C" {

int fednl_cli_main(int argc, char** argv) { return fednl_cli::cli_main(argc, argv, false); }

int fednl_cli_pp_main(int argc, char** argv) { return fednl_cli::cli_main(argc, argv, true); }

int doptExecuteLocal(const char* command, ResultCallbackLocal resultCallback)
{
    std::vector<std::string> parts;
    std::vector<char*> argv = fednl_cli::split_command(command, "bin_fednl_local", parts);
    fednl_cli::CmdLine cl((int)argv.size(), argv.data());
    return fednl_cli::train(cl, resultCallback, nullptr, false);
}

int doptExecuteLocalPP(const char* command, ResultCallbackLocalPP resultCallback)
{
    std::vector<std::string> parts;
    std::vector<char*> argv = fednl_cli::split_command(command, "bin_fednl_local_pp", parts);
    fednl_cli::CmdLine cl((int)argv.size(), argv.data());
    return fednl_cli::train(cl, nullptr, resultCallback, true);
}


int fednl_cli_load_libsvm(const char* path, int add_intercept, int* N, int* d, double* X, double* y)
{
    if (!path || !N || !d)
        return FEDNL_E_MISSING_ARGUMENT;
    fednl_cli::Dataset ds;
    const int rc = fednl_cli::load_libsvm(path, add_intercept != 0, ds, false);
    if (rc != 0)
        return rc;
    *N = ds.N;
    *d = ds.d;
    if (X) {
        if (!y)
            return FEDNL_E_MISSING_ARGUMENT;
        memcpy(X, ds.X.data(), sizeof(double) * ds.X.size());
        memcpy(y, ds.y.data(), sizeof(double) * ds.y.size());
    }
    return 0;
}

uint32_t fednl_cli_crc32(const void* buf, size_t size) { return fednl_cli::crc32_bytes(buf, size); }

}  // extern "C"
