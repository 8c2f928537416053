// Parameter files and command-line overrides, read the way the reference's model and run
// parameters are read (models/parameter-files/lphi4.in; ParameterParser::get as called in
// models/lphi4.h:86-111 and runparameters.h:60-153).
//
// Format: one `key = value [value ...]` per line, `#` starts a comment.  Command-line arguments
// `key=value` override the file; `input=<file>` (or a lone file name) names the file.
#pragma once

#include <array>
#include <cstdint>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace TempLat {

// Marks a parameter as important in the reference (printed in info files); accepted and ignored.
struct ImportantTag {};
constexpr ImportantTag Important{};

class ParameterMissing : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};
class ParameterInvalid : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};

template <typename T> struct ParamConvert;  // specialised per type below

class ParameterParser
{
public:
  ParameterParser() = default;
  ParameterParser(int argc, char **argv)
  {
    std::vector<std::string> overrides;
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      const auto eq = a.find('=');
      if (eq == std::string::npos) {
        readFile(a);
        continue;
      }
      if (a.substr(0, eq) == "input")
        readFile(a.substr(eq + 1));
      else
        overrides.push_back(a);
    }
    for (const auto &o : overrides) setLine(o);
  }

  void readFile(const std::string &fn)
  {
    std::ifstream in(fn);
    if (!in) throw ParameterMissing("cannot open parameter file " + fn);
    std::string line;
    while (std::getline(in, line)) setLine(line);
  }
  // "key = v1 v2 ..." (comments after '#').
  void setLine(std::string line)
  {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    const auto eq = line.find('=');
    if (eq == std::string::npos) return;
    const std::string key = trim(line.substr(0, eq));
    if (key.empty()) return;
    std::istringstream vs(line.substr(eq + 1));
    std::vector<std::string> vals;
    std::string v;
    while (vs >> v) vals.push_back(v);
    values_[key] = vals;
  }
  void set(const std::string &key, const std::string &value) { setLine(key + "=" + value); }
  bool has(const std::string &key) const { return values_.count(key) != 0; }
  void erase(const std::string &key) { values_.erase(key); }

  template <typename T> T get(const std::string &key, ImportantTag = {}) const
  {
    return ParamConvert<T>::from(key, one(key));
  }
  template <typename T, typename D, typename = std::enable_if_t<!std::is_same_v<D, ImportantTag>>>
  T get(const std::string &key, const D &def, ImportantTag = {}) const
  {
    if (!has(key)) return T(def);
    return ParamConvert<T>::from(key, one(key));
  }
  template <typename T, size_t N> std::array<T, N> get(const std::string &key, ImportantTag = {}) const
  {
    return many<T, N>(key);
  }
  template <typename T, size_t N> std::array<T, N> get(const std::string &key, const std::array<T, N> &def,
                                                       ImportantTag = {}) const
  {
    if (!has(key)) return def;
    return many<T, N>(key);
  }
  template <typename T, size_t N> std::array<T, N> get(const std::string &key, T def, ImportantTag = {}) const
  {
    if (!has(key)) {
      std::array<T, N> a;
      a.fill(def);
      return a;
    }
    return many<T, N>(key);
  }
  template <typename T, typename... A> T getOverride(const std::string &key, A &&...a) const
  {
    return get<T>(key, std::forward<A>(a)...);
  }
  // Seed string of the random initial conditions (the reference keys its generator by a string).
  std::string getSeed(const std::string &key) const { return has(key) ? joined(key) : std::string("1"); }

  const std::map<std::string, std::vector<std::string>> &all() const { return values_; }

private:
  static std::string trim(const std::string &s)
  {
    const auto b = s.find_first_not_of(" \t\r");
    const auto e = s.find_last_not_of(" \t\r");
    return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
  }
  const std::string &one(const std::string &key) const
  {
    auto it = values_.find(key);
    if (it == values_.end() || it->second.empty()) throw ParameterMissing("parameter '" + key + "' is missing");
    return it->second[0];
  }
  std::string joined(const std::string &key) const
  {
    std::string r;
    for (const auto &v : values_.at(key)) r += (r.empty() ? "" : " ") + v;
    return r;
  }
  template <typename T, size_t N> std::array<T, N> many(const std::string &key) const
  {
    auto it = values_.find(key);
    if (it == values_.end()) throw ParameterMissing("parameter '" + key + "' is missing");
    if (it->second.size() != N)
      throw ParameterInvalid("parameter '" + key + "' needs " + std::to_string(N) + " values, got " +
                             std::to_string(it->second.size()));
    std::array<T, N> a;
    for (size_t i = 0; i < N; ++i) a[i] = ParamConvert<T>::from(key, it->second[i]);
    return a;
  }
  std::map<std::string, std::vector<std::string>> values_;
};

template <> struct ParamConvert<double> {
  static double from(const std::string &k, const std::string &s)
  {
    try {
      size_t n = 0;
      double v = std::stod(s, &n);
      if (n != s.size()) throw std::invalid_argument(s);
      return v;
    } catch (const std::exception &) {
      throw ParameterInvalid("parameter '" + k + "': '" + s + "' is not a number");
    }
  }
};
template <> struct ParamConvert<float> {
  static float from(const std::string &k, const std::string &s) { return (float)ParamConvert<double>::from(k, s); }
};
template <> struct ParamConvert<long long> {
  static long long from(const std::string &k, const std::string &s)
  {
    try {
      size_t n = 0;
      long long v = std::stoll(s, &n);
      if (n != s.size()) throw std::invalid_argument(s);
      return v;
    } catch (const std::exception &) {
      throw ParameterInvalid("parameter '" + k + "': '" + s + "' is not an integer");
    }
  }
};
template <> struct ParamConvert<long> {
  static long from(const std::string &k, const std::string &s) { return (long)ParamConvert<long long>::from(k, s); }
};
template <> struct ParamConvert<int> {
  static int from(const std::string &k, const std::string &s) { return (int)ParamConvert<long long>::from(k, s); }
};
template <> struct ParamConvert<unsigned long> {
  static unsigned long from(const std::string &k, const std::string &s)
  {
    const long long v = ParamConvert<long long>::from(k, s);
    if (v < 0) throw ParameterInvalid("parameter '" + k + "' must be non-negative");
    return (unsigned long)v;
  }
};
template <> struct ParamConvert<bool> {
  static bool from(const std::string &k, const std::string &s)
  {
    if (s == "true" || s == "1" || s == "True" || s == "yes") return true;
    if (s == "false" || s == "0" || s == "False" || s == "no") return false;
    throw ParameterInvalid("parameter '" + k + "': '" + s + "' is not a boolean");
  }
};
template <> struct ParamConvert<std::string> {
  static std::string from(const std::string &, const std::string &s) { return s; }
};

}  // namespace TempLat
