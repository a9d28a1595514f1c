// logging.cpp -- ROCBLAS_LAYER trace / bench / profile logs of the trsm
// entry points (formats and sources in logging.hpp).
//
// A log file is opened (truncated) once per process and shared by every
// handle that names it; every line is written whole under that file's lock.
// The profile keeps one count per distinct argument set across all handles
// and writes the list when the process exits (also on quick_exit).
#include "logging.hpp"

#include <cstdio>
#include <cstdlib>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace rbamd
{
struct LogSink
{
    std::mutex mu;
    FILE*      f = nullptr;

    void write(const std::string& s)
    {
        std::lock_guard<std::mutex> g(mu);
        std::fwrite(s.data(), 1, s.size(), f);
        std::fflush(f);
    }
};

namespace
{
// Sinks live until the process ends (never destroyed), so the profile can
// still be written from a static destructor.
std::mutex                                      g_sinks_mu;
std::map<std::string, std::unique_ptr<LogSink>>* g_sinks = nullptr;

LogSink* open_sink(const char* env_name)
{
    const char* path = std::getenv(env_name);
    if(!path)
        path = std::getenv("ROCBLAS_LOG_PATH");
    const std::string           key = path ? path : "";
    std::lock_guard<std::mutex> g(g_sinks_mu);
    if(!g_sinks)
        g_sinks = new std::map<std::string, std::unique_ptr<LogSink>>;
    auto it = g_sinks->find(key);
    if(it != g_sinks->end())
        return it->second.get();
    auto sink = std::make_unique<LogSink>();
    sink->f   = stderr;
    if(path)
    {
        sink->f = std::fopen(path, "w");
        if(!sink->f)
        {
            std::fprintf(stderr, "rocblas_amd: cannot open log file %s, logging to stderr\n",
                         path);
            sink->f = stderr;
        }
    }
    LogSink* raw = sink.get();
    g_sinks->emplace(key, std::move(sink));
    return raw;
}

// Distinct argument sets in first-seen order, with their call counts.
struct Profile
{
    std::mutex                              mu;
    LogSink*                                sink = nullptr;
    std::vector<std::string>                keys;
    std::unordered_map<std::string, size_t> counts;

    void add(LogSink* s, std::string key)
    {
        std::lock_guard<std::mutex> g(mu);
        if(!sink)
            sink = s;
        auto it = counts.find(key);
        if(it != counts.end())
        {
            ++it->second;
            return;
        }
        counts.emplace(key, 1);
        keys.push_back(std::move(key));
    }

    void dump()
    {
        std::lock_guard<std::mutex> g(mu);
        if(!sink || keys.empty())
            return;
        std::string out;
        for(const auto& k : keys)
            out += "- { " + k + ", call_count: " + std::to_string(counts[k]) + " }\n";
        sink->write(out);
        keys.clear();
        counts.clear();
    }

    ~Profile()
    {
        dump();
    }
};

Profile& profile()
{
    static Profile p;
    static int     registered = std::at_quick_exit([] { profile().dump(); });
    (void)registered;
    return p;
}

char side_letter(rocblas_side s)
{
    switch(s)
    {
    case rocblas_side_left: return 'L';
    case rocblas_side_right: return 'R';
    case rocblas_side_both: return 'B';
    }
    return ' ';
}

char fill_letter(rocblas_fill f)
{
    switch(f)
    {
    case rocblas_fill_upper: return 'U';
    case rocblas_fill_lower: return 'L';
    case rocblas_fill_full: return 'F';
    }
    return ' ';
}

char op_letter(rocblas_operation t)
{
    switch(t)
    {
    case rocblas_operation_none: return 'N';
    case rocblas_operation_transpose: return 'T';
    case rocblas_operation_conjugate_transpose: return 'C';
    }
    return ' ';
}

char diag_letter(rocblas_diagonal d)
{
    switch(d)
    {
    case rocblas_diagonal_non_unit: return 'N';
    case rocblas_diagonal_unit: return 'U';
    }
    return ' ';
}

template <typename T>
constexpr char prec_letter();
template <>
constexpr char prec_letter<float>()
{
    return 's';
}
template <>
constexpr char prec_letter<double>()
{
    return 'd';
}

template <typename T>
constexpr const char* prec_string();
template <>
constexpr const char* prec_string<float>()
{
    return "f32_r";
}
template <>
constexpr const char* prec_string<double>()
{
    return "f64_r";
}

// alpha as the logs print it: NaN for a null pointer, read from the device in
// device pointer mode
template <typename T>
double alpha_value(rocblas_handle h, const T* alpha)
{
    if(!alpha)
        return std::numeric_limits<double>::quiet_NaN();
    if(h->pointer_mode == rocblas_pointer_mode_host)
        return double(*alpha);
    T host = T(0);
    if(hipMemcpyAsync(&host, alpha, sizeof(T), hipMemcpyDeviceToHost, h->stream) != hipSuccess
       || hipStreamSynchronize(h->stream) != hipSuccess)
        return std::numeric_limits<double>::quiet_NaN();
    return double(host);
}

// a double the way the reference's YAML-free streams print it (ostream
// default: six significant digits)
std::string num(double x)
{
    std::ostringstream os;
    os << x;
    return os.str();
}

std::string ptr(const void* p)
{
    std::ostringstream os;
    os << p;
    return os.str();
}

} // namespace

void init_logging(rocblas_handle h)
{
    const char* mode = std::getenv("ROCBLAS_LAYER");
    if(!mode)
        return;
    h->layer_mode = int(std::strtol(mode, nullptr, 0));
    if(h->layer_mode & 1)
        h->log_trace_os = open_sink("ROCBLAS_LOG_TRACE_PATH");
    if(h->layer_mode & 2)
        h->log_bench_os = open_sink("ROCBLAS_LOG_BENCH_PATH");
    if(h->layer_mode & 4)
        h->log_profile_os = open_sink("ROCBLAS_LOG_PROFILE_PATH");
}

template <typename T>
void log_trsm(rocblas_handle    h,
              TrsmForm          form,
              bool              api64,
              rocblas_side      side,
              rocblas_fill      uplo,
              rocblas_operation transA,
              rocblas_diagonal  diag,
              int64_t           m,
              int64_t           n,
              const T*          alpha,
              const void*       A,
              int64_t           lda,
              rocblas_stride    stride_a,
              const void*       B,
              int64_t           ldb,
              rocblas_stride    stride_b,
              int64_t           batch_count)
{
    if(!h || !(h->layer_mode & 7) || h->size_query)
        return;
    const char sL = side_letter(side), uL = fill_letter(uplo), tL = op_letter(transA),
               dL = diag_letter(diag);
    // function names as the reference spells them (rocblas_trsm*_imp.hpp:41-52)
    std::string name = form == TrsmForm::plain     ? "rocblas_"
                       : form == TrsmForm::batched ? "rocblas_batched_"
                                                   : "rocblas_strided_batched_";
    name += prec_letter<T>();
    name += "trsm";
    if(api64)
        name += "_64";
    const double a = alpha_value(h, alpha);

    if(h->layer_mode & 1)
    {
        std::string s = name + "," + sL + "," + uL + "," + tL + "," + dL + ","
                        + std::to_string(m) + "," + std::to_string(n) + "," + num(a) + ","
                        + ptr(A) + "," + std::to_string(lda) + ",";
        if(form == TrsmForm::strided)
            s += std::to_string(stride_a) + ",";
        s += ptr(B) + "," + std::to_string(ldb);
        if(form == TrsmForm::strided)
            s += "," + std::to_string(stride_b);
        if(form != TrsmForm::plain)
            s += "," + std::to_string(batch_count);
        s += ",atomics_allowed\n";
        h->log_trace_os->write(s);
    }
    if(h->layer_mode & 2)
    {
        std::string s = api64 ? "./rocblas-bench --api 1" : "./rocblas-bench";
        s += form == TrsmForm::plain     ? " -f trsm"
             : form == TrsmForm::batched ? " -f trsm_batched"
                                         : " -f trsm_strided_batched";
        s += std::string(" -r ") + prec_string<T>() + " --side " + sL + " --uplo " + uL
             + " --transposeA " + tL + " --diag " + dL + " -m " + std::to_string(m) + " -n "
             + std::to_string(n) + " --alpha " + num(a) + " --lda " + std::to_string(lda);
        if(form == TrsmForm::strided)
            s += " --stride_a " + std::to_string(stride_a);
        s += " --ldb " + std::to_string(ldb);
        if(form == TrsmForm::strided)
            s += " --stride_b " + std::to_string(stride_b);
        if(form != TrsmForm::plain)
            s += " --batch_count " + std::to_string(batch_count);
        s += "\n";
        h->log_bench_os->write(s);
    }
    if(h->layer_mode & 4)
    {
        auto q = [](char c) { return std::string("'") + c + "'"; };
        std::string s = "rocblas_function: \"" + name + "\", atomics_mode: atomics_allowed, side: "
                        + q(sL) + ", uplo: " + q(uL) + ", transA: " + q(tL) + ", diag: " + q(dL)
                        + ", m: " + std::to_string(m) + ", n: " + std::to_string(n)
                        + ", lda: " + std::to_string(lda);
        if(form == TrsmForm::strided)
            s += ", stride_A: " + std::to_string(stride_a);
        s += ", ldb: " + std::to_string(ldb);
        if(form == TrsmForm::strided)
            s += ", stride_B: " + std::to_string(stride_b);
        if(form != TrsmForm::plain)
            s += ", batch_count: " + std::to_string(batch_count);
        profile().add(h->log_profile_os, std::move(s));
    }
}

#define RBAMD_INSTANTIATE_LOG(T_)                                                                \
    template void log_trsm<T_>(rocblas_handle, TrsmForm, bool, rocblas_side, rocblas_fill,       \
                               rocblas_operation, rocblas_diagonal, int64_t, int64_t, const T_*, \
                               const void*, int64_t, rocblas_stride, const void*, int64_t,       \
                               rocblas_stride, int64_t);
RBAMD_INSTANTIATE_LOG(float)
RBAMD_INSTANTIATE_LOG(double)
#undef RBAMD_INSTANTIATE_LOG

} // namespace rbamd
