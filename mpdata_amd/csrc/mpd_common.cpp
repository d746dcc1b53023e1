#include "mpd_common.hpp"

#include <dlfcn.h>

#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

namespace mpd {

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

int hip_fail(hipError_t e, const char* what)
{
    return fail(e == hipErrorOutOfMemory ? MPD_ERR_NOMEM : MPD_ERR_DEVICE, "%s failed: %s",
                what, hipGetErrorString(e));
}

int ensure_device()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(MPD_ERR_DEVICE, "no HIP device available (%s)",
                    e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    return MPD_OK;
}

static std::mutex g_stream_mu;
static std::vector<hipStream_t> g_streams;

hipStream_t stream()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if ((int)g_streams.size() <= dev) g_streams.resize(dev + 1, nullptr);
    if (!g_streams[dev]) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        g_streams[dev] = s;
    }
    return g_streams[dev];
}

int DevBuf::alloc(size_t count)
{
    if (p && n == count) return MPD_OK;
    release();
    if (count == 0) return MPD_OK;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(double));
    if (e != hipSuccess) {
        p = nullptr;
        return fail(MPD_ERR_NOMEM, "hipMalloc of %zu doubles failed: %s", count,
                    hipGetErrorString(e));
    }
    n = count;
    return MPD_OK;
}

static std::mutex g_tune_mu;
static std::map<std::string, std::vector<int>> g_tune;

// The table shipped beside the library (tuning.txt in the directory of this shared
// object): results measured on MI355X for the shapes of the benchmark configurations.
static std::string shipped_table()
{
    Dl_info info;
    if (!dladdr((void*)&shipped_table, &info) || !info.dli_fname) return "";
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tuning.txt";
}

bool tune_lookup(const std::string& key, int n, int* v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it == g_tune.end()) {
        const char* env = getenv("MPD_TUNE_CACHE");
        const std::string path = env ? std::string(env) : shipped_table();
        FILE* f = path.empty() ? nullptr : fopen(path.c_str(), "r");
        if (!f) return false;
        char line[1 << 16];
        while (fgets(line, sizeof line, f)) {
            std::istringstream in(line);
            std::string k;
            in >> k;
            if (k != key) continue;
            std::vector<int> vals(n);
            bool ok = true;
            for (int i = 0; i < n && ok; i++) ok = bool(in >> vals[i]);
            if (ok) g_tune[key] = vals;
        }
        fclose(f);
        it = g_tune.find(key);
        if (it == g_tune.end()) return false;
    }
    if ((int)it->second.size() != n) return false;
    for (int i = 0; i < n; i++) v[i] = it->second[i];
    return true;
}

void tune_store(const std::string& key, int n, const int* v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = std::vector<int>(v, v + n);
    const char* path = getenv("MPD_TUNE_CACHE");
    if (FILE* f = path ? fopen(path, "a") : nullptr) {
        fprintf(f, "%s", key.c_str());
        for (int i = 0; i < n; i++) fprintf(f, " %d", v[i]);
        fprintf(f, "\n");
        fclose(f);
    }
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

}  // namespace mpd

extern "C" const char* mpd_error_string(void) { return mpd::g_err.c_str(); }

extern "C" int mpd_device_count(int* n)
{
    MPD_API_BEGIN
    if (!n) return mpd::fail(MPD_ERR_ARG, "mpd_device_count: NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return MPD_OK;
    MPD_API_END
}
