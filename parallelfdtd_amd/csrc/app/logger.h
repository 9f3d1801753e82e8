// logger.h -- the logging interface the reference's callers use (src/logger.h:33-85):
// loggerInit() and log_msg<LEVEL>(L"format") % arg % arg, without Boost.
//
// The format string takes printf-style directives (%d, %u, %f, %.3f, %s, %e, %x, ... and %%),
// replaced by the `%`-chained arguments in order when the message ends (the end of the
// full expression). As in the reference, messages at LOG_ERROR and more severe go to the wide
// standard output, and messages at LOG_INFO and more severe are appended to solver_log.txt in
// the working directory (loggerInit truncates it, src/logger.cpp).
#pragma once

#include <cstdio>
#include <cwchar>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#define LOG_COUT 2
#define LOG_TO_FILE 4

enum log_level { LOG_NOTHING, LOG_CRITICAL, LOG_ERROR, LOG_WARNING, LOG_INFO, LOG_DEBUG, LOG_VERBOSE };

class Logger {
 public:
  Logger(log_level level, const wchar_t *msg) : level_(level), fmt_(msg ? msg : L"") {}
  // Only the last holder of a message emits it.
  Logger(const Logger &o) : level_(o.level_), fmt_(o.fmt_), args_(o.args_) { o.live_ = false; }
  Logger &operator=(const Logger &) = delete;
  ~Logger() {
    if (!live_) return;
    const std::wstring line = text();
    if (LOG_COUT >= level_) std::wcout << (int)level_ << L" " << line << std::endl;
    if (LOG_TO_FILE >= level_) {
      std::wofstream f("solver_log.txt", std::fstream::out | std::fstream::app);
      f << (int)level_ << L" " << line << std::endl;
    }
  }

  template <typename T>
  Logger &operator%(const T &value) {
    Arg a;
    if constexpr (std::is_floating_point<T>::value) {
      a.kind = Arg::kFloat;
      a.f = (double)value;
    } else if constexpr (std::is_integral<T>::value && std::is_signed<T>::value) {
      a.kind = Arg::kSigned;
      a.i = (long long)value;
    } else if constexpr (std::is_integral<T>::value || std::is_enum<T>::value) {
      a.kind = Arg::kUnsigned;
      a.u = (unsigned long long)value;
    } else {
      std::wostringstream s;
      put(s, value);
      a.kind = Arg::kText;
      a.s = s.str();
    }
    args_.push_back(a);
    return *this;
  }

  // The message with every directive replaced (what the destructor writes).
  std::wstring text() const {
    std::wstring out;
    size_t next = 0;
    for (size_t i = 0; i < fmt_.size(); ++i) {
      if (fmt_[i] != L'%') {
        out += fmt_[i];
        continue;
      }
      if (i + 1 < fmt_.size() && fmt_[i + 1] == L'%') {
        out += L'%';
        ++i;
        continue;
      }
      size_t j = i + 1;  // flags, width, precision, length, conversion
      while (j < fmt_.size() && std::wcschr(L"-+ #0123456789.hlLqjzt", fmt_[j])) ++j;
      if (j >= fmt_.size()) {
        out += fmt_.substr(i);
        break;
      }
      std::wstring spec = fmt_.substr(i, j - i);  // '%' + flags/width/precision (+length)
      while (!spec.empty() && std::wcschr(L"hlLqjzt", spec.back())) spec.pop_back();
      out += next < args_.size() ? args_[next].format(spec, fmt_[j]) : std::wstring();
      ++next;
      i = j;
    }
    return out;
  }

 private:
  struct Arg {
    enum Kind { kSigned, kUnsigned, kFloat, kText } kind = kText;
    long long i = 0;
    unsigned long long u = 0;
    double f = 0.0;
    std::wstring s;
    std::wstring format(const std::wstring &spec, wchar_t conv) const {
      wchar_t buf[512];
      std::wstring f2 = spec;
      const bool fconv = std::wcschr(L"fFeEgGaA", conv) != nullptr;
      const bool iconv = std::wcschr(L"diouxXc", conv) != nullptr;
      if (kind == kFloat && (fconv || iconv)) {
        f2 += fconv ? conv : L'f';
        std::swprintf(buf, 512, f2.c_str(), f);
      } else if ((kind == kSigned || kind == kUnsigned) && fconv) {
        f2 += conv;
        std::swprintf(buf, 512, f2.c_str(), kind == kSigned ? (double)i : (double)u);
      } else if (kind == kSigned && iconv && conv != L'c') {
        f2 += L"ll";
        f2 += (conv == L'd' || conv == L'i') ? conv : L'd';
        std::swprintf(buf, 512, f2.c_str(), i);
      } else if (kind == kUnsigned && iconv && conv != L'c') {
        f2 += L"ll";
        f2 += (conv == L'd' || conv == L'i') ? L'u' : conv;
        std::swprintf(buf, 512, f2.c_str(), u);
      } else {
        std::wostringstream o;
        if (kind == kSigned) o << i;
        else if (kind == kUnsigned) o << u;
        else if (kind == kFloat) o << f;
        else o << s;
        return o.str();
      }
      return buf;
    }
  };

  template <typename T>
  static void put(std::wostringstream &s, const T &v) {
    s << v;
  }
  static void put(std::wostringstream &s, const std::string &v) {
    s << std::wstring(v.begin(), v.end());
  }
  static void put(std::wostringstream &s, const char *v) { s << (v ? v : "(null)"); }

  log_level level_;
  std::wstring fmt_;
  std::vector<Arg> args_;
  mutable bool live_ = true;
};

template <log_level level>
Logger log_msg(const wchar_t *msg) {
  return Logger(level, msg);
}

// Truncates solver_log.txt (src/logger.cpp). Defined in libpfdtd_app.so.
void loggerInit();
