// tune_cache.h -- the batched kernels' autotuner decisions, kept in memory and
// in a text file across processes (host code only).
//
// File format: one decision per line, "<key> <shape>", where key is
// "<device>/<kernel>/<rows class>/<width class>" (no spaces) and shape an
// integer code; lines that do not parse are skipped.  New decisions are
// appended.  Path: hz_batch_tune_cache(), else the HZ_TUNE_CACHE environment
// variable ("" or "off": memory only), else
// ${XDG_CACHE_HOME:-$HOME/.cache}/hazmath_mi355x/tune.txt.
#pragma once

#include <map>
#include <mutex>
#include <string>

namespace hz {

class TuneCache {
  public:
    static TuneCache &get();
    // Replaces the in-memory decisions with those of the file at path (the
    // default path when path is null and use_default is set; memory only when
    // path is null otherwise).  Returns the number of decisions read.
    int reset(const char *path, bool use_default);
    int lookup(const std::string &key);   // -1 when absent
    void store(const std::string &key, int shape);
    std::string path();

  private:
    void ensure_init();
    int load_file();
    std::map<std::string, int> m_;
    std::string path_;
    bool init_ = false;
    std::mutex mu_;
};

// Parses one cache line; false when it is not "<key> <integer>".
bool parse_tune_line(const std::string &line, std::string &key, int &shape);

}  // namespace hz
