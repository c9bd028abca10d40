// Autotuner cache: configurations chosen by timing, kept per process and,
// when a cache file is named (bk_tune_set_cache or the BBRCIT_HIP_TUNE_CACHE
// environment variable), saved to that file and reloaded from it.
//
// File format (text, one entry per line):
//     bbrcit-hip-tune <format version> <tuning version>
//     <key> TAB <value> <value> ...
// A file whose header names another format or tuning version is ignored (its
// entries belong to other kernels) and rewritten on the next save.  Keys name
// the device (architecture and compute-unit count) and the workload shape;
// values are small integers whose meaning the caller defines.  Saves write a
// temporary file next to the cache and rename it over the cache, so a reader
// never sees a partial file.
#include "../../include/bbrcit_hip.h"
#include "bk_common.h"
#include "bk_internal.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

namespace bk {

namespace {

constexpr int kTuneFormat = 1;

struct TuneStore {
  std::mutex mu;
  bool env_read = false;
  std::string path;  // empty: memory only
  std::map<std::string, std::vector<int>> entries;
};

TuneStore &store() {
  static TuneStore s;
  return s;
}

std::string header() {
  return "bbrcit-hip-tune " + std::to_string(kTuneFormat) + " " + std::to_string(kTuneVersion);
}

// Entries of the file at `path` (none when it is missing, unreadable or of
// another version).
std::map<std::string, std::vector<int>> read_file(const std::string &path) {
  std::map<std::string, std::vector<int>> out;
  std::ifstream in(path);
  if (!in) return out;
  std::string line;
  if (!std::getline(in, line) || line != header()) return out;
  while (std::getline(in, line)) {
    const size_t tab = line.find('\t');
    if (tab == std::string::npos || tab == 0) continue;
    std::istringstream vs(line.substr(tab + 1));
    std::vector<int> v;
    int x;
    while (vs >> x) v.push_back(x);
    if (!v.empty()) out[line.substr(0, tab)] = v;
  }
  return out;
}

void write_file(const std::string &path, const std::map<std::string, std::vector<int>> &entries) {
  const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (!out) fail(BK_ERR_INVALID_ARG, "cannot write the tuning cache " + tmp);
    out << header() << "\n";
    for (const auto &e : entries) {
      out << e.first << "\t";
      for (size_t i = 0; i < e.second.size(); ++i) out << (i ? " " : "") << e.second[i];
      out << "\n";
    }
    if (!out) fail(BK_ERR_INVALID_ARG, "cannot write the tuning cache " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) fail(BK_ERR_INVALID_ARG, "cannot replace the tuning cache " + path);
}

// Caller holds the lock.
void read_env_once(TuneStore &s) {
  if (s.env_read) return;
  s.env_read = true;
  const char *p = std::getenv("BBRCIT_HIP_TUNE_CACHE");
  if (p && *p) {
    s.path = p;
    for (auto &e : read_file(s.path)) s.entries.insert(e);
  }
}

}  // namespace

std::string device_tune_tag() {
  static std::mutex mu;
  static std::map<int, std::string> tags;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(mu);
  const auto it = tags.find(dev);
  if (it != tags.end()) return it->second;
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  std::string arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch = arch.substr(0, colon);
  return tags[dev] = arch + "/" + std::to_string(prop.multiProcessorCount) + "cu";
}

bool tune_lookup(const std::string &key, std::vector<int> &value) {
  TuneStore &s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  read_env_once(s);
  const auto it = s.entries.find(key);
  if (it == s.entries.end()) return false;
  value = it->second;
  return true;
}

void tune_record(const std::string &key, const std::vector<int> &value) {
  if (key.find('\t') != std::string::npos || key.find('\n') != std::string::npos)
    fail(BK_ERR_INVALID_ARG, "internal: bad tuning key");
  TuneStore &s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  read_env_once(s);
  s.entries[key] = value;
  if (s.path.empty()) return;
  // merge with what other processes may have saved meanwhile, then save
  auto merged = read_file(s.path);
  for (const auto &e : s.entries) merged[e.first] = e.second;
  write_file(s.path, merged);
}

}  // namespace bk

extern "C" {

int bk_tune_set_cache(const char *path) {
  return bk::guarded([&] {
    bk::TuneStore &s = bk::store();
    std::lock_guard<std::mutex> lock(s.mu);
    s.env_read = true;  // an explicit choice overrides the environment
    s.entries.clear();
    s.path = path ? path : "";
    if (!s.path.empty())
      for (auto &e : bk::read_file(s.path)) s.entries.insert(e);
  });
}

int bk_tune_entry_count(size_t *count_out) {
  return bk::guarded([&] {
    if (!count_out) bk::fail(BK_ERR_INVALID_ARG, "null output pointer");
    bk::TuneStore &s = bk::store();
    std::lock_guard<std::mutex> lock(s.mu);
    bk::read_env_once(s);
    *count_out = s.entries.size();
  });
}

}  // extern "C"
