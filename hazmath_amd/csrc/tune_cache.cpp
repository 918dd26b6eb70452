// tune_cache.cpp -- on-disk cache of the batched kernels' autotuner decisions.
#include "tune_cache.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace hz {

namespace {

std::string default_path() {
    const char *env = std::getenv("HZ_TUNE_CACHE");
    if (env) {
        if (!*env || std::strcmp(env, "off") == 0) return std::string();
        return env;
    }
    std::string dir;
    const char *xdg = std::getenv("XDG_CACHE_HOME");
    const char *home = std::getenv("HOME");
    if (xdg && *xdg) dir = xdg;
    else if (home && *home) dir = std::string(home) + "/.cache";
    else return std::string();
    return dir + "/hazmath_mi355x/tune.txt";
}

// mkdir -p of the directory part of path; false on failure
bool make_parent_dirs(const std::string &path) {
    for (size_t pos = path.find('/', 1); pos != std::string::npos; pos = path.find('/', pos + 1)) {
        const std::string d = path.substr(0, pos);
        if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    return true;
}

}  // namespace

bool parse_tune_line(const std::string &line, std::string &key, int &shape) {
    std::istringstream is(line);
    std::string k, v, extra;
    if (!(is >> k >> v) || (is >> extra)) return false;
    if (k.empty() || k[0] == '#') return false;
    char *end = nullptr;
    errno = 0;
    const long s = std::strtol(v.c_str(), &end, 10);
    if (errno != 0 || end == v.c_str() || *end != '\0' || s <= 0 || s > 1000000) return false;
    key = k;
    shape = (int)s;
    return true;
}

TuneCache &TuneCache::get() {
    static TuneCache c;
    return c;
}

int TuneCache::load_file() {
    if (path_.empty()) return 0;
    std::ifstream f(path_);
    if (!f) return 0;
    int count = 0;
    std::string line, key;
    int shape = 0;
    while (std::getline(f, line))
        if (parse_tune_line(line, key, shape)) {
            m_[key] = shape;
            ++count;
        }
    return count;
}

void TuneCache::ensure_init() {
    if (init_) return;
    init_ = true;
    path_ = default_path();
    load_file();
}

int TuneCache::reset(const char *path, bool use_default) {
    std::lock_guard<std::mutex> g(mu_);
    m_.clear();
    init_ = true;
    path_ = path ? std::string(path) : (use_default ? default_path() : std::string());
    return load_file();
}

int TuneCache::lookup(const std::string &key) {
    std::lock_guard<std::mutex> g(mu_);
    ensure_init();
    auto it = m_.find(key);
    return it == m_.end() ? -1 : it->second;
}

void TuneCache::store(const std::string &key, int shape) {
    std::lock_guard<std::mutex> g(mu_);
    ensure_init();
    m_[key] = shape;
    if (path_.empty() || !make_parent_dirs(path_)) return;
    // append; a cache that cannot be written only costs a re-tune next time
    if (FILE *f = std::fopen(path_.c_str(), "a")) {
        std::fprintf(f, "%s %d\n", key.c_str(), shape);
        std::fclose(f);
    }
}

std::string TuneCache::path() {
    std::lock_guard<std::mutex> g(mu_);
    ensure_init();
    return path_;
}

}  // namespace hz
