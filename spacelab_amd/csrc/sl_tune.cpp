// Persistent cache of the autotuned launch choices.
//
// The fast engine's tile split (wavefronts per 64x64 tile) sets the order of the partial
// sums, so two runs that tune the same configuration separately can pick different splits
// when two choices time alike, and then differ in the last bits.  The choice made for a
// configuration is therefore stored and read back: once a configuration has been tuned
// on a machine, every later run of it uses the same split and repeats bit for bit.  The
// ordered pass's launch shape is stored the same way (there only the time depends on it).
//
// Two files, both a flat JSON object {"<key>": {"split": s, "ms": [...]}, ...} (for the
// ordered pass "split" holds 1 + the shape index):
//   * the user cache, read and written: $SPACELAB_TUNE_CACHE, else
//     $XDG_CACHE_HOME/spacelab_amd/tune.json, else $HOME/.cache/spacelab_amd/tune.json
//     (SPACELAB_TUNE_CACHE=off disables reading and writing it);
//   * the shipped cache, read only: data/tune.json beside the library's lib/ folder
//     (choices measured on MI355X for the benchmark configurations).
// The user cache wins over the shipped one.
#include "sl_internal.hpp"

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

namespace sl {
namespace {

struct Entry {
    int split = 0;
    std::string ms;   // the timings as stored (kept verbatim when the file is rewritten)
};

std::mutex g_mu;
bool g_loaded = false;
std::map<std::string, Entry> g_user, g_shipped;

std::string user_path()
{
    const char* e = std::getenv("SPACELAB_TUNE_CACHE");
    if (e && *e) return (std::strcmp(e, "off") == 0 || std::strcmp(e, "0") == 0) ? std::string() : std::string(e);
    if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) return std::string(x) + "/spacelab_amd/tune.json";
    if (const char* h = std::getenv("HOME"); h && *h) return std::string(h) + "/.cache/spacelab_amd/tune.json";
    return std::string();
}

std::string shipped_path()
{
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&shipped_path), &info) == 0 || !info.dli_fname) return std::string();
    std::string p(info.dli_fname);
    const size_t s = p.rfind('/');
    if (s == std::string::npos) return std::string();
    p.resize(s);                                   // .../spacelab_amd/lib
    const size_t s2 = p.rfind('/');
    if (s2 == std::string::npos) return std::string();
    p.resize(s2);                                  // .../spacelab_amd
    return p + "/data/tune.json";
}

// Reads entries of the form "key": {"split": N, "ms": [...]} (the layout write() makes).
void parse(const std::string& path, std::map<std::string, Entry>& out)
{
    if (path.empty()) return;
    std::ifstream f(path);
    if (!f) return;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string t = ss.str();
    size_t pos = 0;
    while (true) {
        const size_t k0 = t.find('"', pos);
        if (k0 == std::string::npos) break;
        const size_t k1 = t.find('"', k0 + 1);
        if (k1 == std::string::npos) break;
        const std::string key = t.substr(k0 + 1, k1 - k0 - 1);
        const size_t b0 = t.find('{', k1);
        const size_t b1 = b0 == std::string::npos ? b0 : t.find('}', b0);
        if (b1 == std::string::npos) break;
        const std::string body = t.substr(b0, b1 - b0 + 1);
        const size_t sp = body.find("\"split\":");
        if (sp != std::string::npos) {
            Entry e;
            e.split = std::atoi(body.c_str() + sp + 8);
            const size_t m0 = body.find('['), m1 = body.find(']');
            if (m0 != std::string::npos && m1 != std::string::npos && m1 > m0) e.ms = body.substr(m0, m1 - m0 + 1);
            if (e.split >= 1 && e.split <= 64) out[key] = e;
        }
        pos = b1 + 1;
    }
}

void load()
{
    if (g_loaded) return;
    g_loaded = true;
    parse(shipped_path(), g_shipped);
    parse(user_path(), g_user);
}

void make_dirs(const std::string& file)
{
    for (size_t p = file.find('/', 1); p != std::string::npos; p = file.find('/', p + 1)) {
        const std::string d = file.substr(0, p);
        if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return;
    }
}

void write_user()
{
    const std::string path = user_path();
    if (path.empty()) return;
    make_dirs(path);
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream f(tmp);
        if (!f) return;
        f << "{\n";
        size_t k = 0;
        for (const auto& kv : g_user) {
            f << "  \"" << kv.first << "\": {\"split\": " << kv.second.split;
            if (!kv.second.ms.empty()) f << ", \"ms\": " << kv.second.ms;
            f << "}" << (++k < g_user.size() ? ",\n" : "\n");
        }
        f << "}\n";
        if (!f) return;
    }
    std::rename(tmp.c_str(), path.c_str());
}

}  // namespace

std::string tune_key(int precision, int gravity, int n_sys, int n_used, int nb_max, int n_tiles)
{
    char b[160];
    std::snprintf(b, sizeof b, "tiles:prec%d:grav%d:sys%d:slots%d:nbmax%d:tiles%d", precision, gravity, n_sys,
                  n_used, nb_max, n_tiles);
    return std::string(b);
}

int tune_cache_get(const std::string& key)
{
    std::lock_guard<std::mutex> lk(g_mu);
    load();
    auto it = g_user.find(key);
    if (it != g_user.end()) return it->second.split;
    it = g_shipped.find(key);
    return it != g_shipped.end() ? it->second.split : 0;
}

void tune_cache_put(const std::string& key, int split, const float* ms, int n)
{
    std::lock_guard<std::mutex> lk(g_mu);
    load();
    Entry e;
    e.split = split;
    std::ostringstream os;
    os << "[";
    for (int k = 0; k < n; k++) os << (k ? ", " : "") << ms[k];
    os << "]";
    e.ms = os.str();
    g_user[key] = e;
    write_user();
}

}  // namespace sl
