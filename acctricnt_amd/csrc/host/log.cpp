#include "log.h"

#include <chrono>
#include <cstdarg>
#include <cstring>
#include <ctime>
#include <locale>
#include <mutex>
#include <sstream>

namespace {
std::mutex g_mu;
FILE *g_fp = nullptr;
const auto g_start = std::chrono::steady_clock::now();
const char *kNames[] = {"TRACE", "DEBUG", "INFO", "WARN", "ERROR", "FATAL"};

void emit(FILE *f, const char *stamp, int level, double ts, double et, const char *file, int line,
          const char *fmt, va_list ap) {
  std::fprintf(f, "%s %-5s (ts: %.6lf s, et: %.6lf s) %s:%d: ", stamp, kNames[level], ts, et, file, line);
  std::vfprintf(f, fmt, ap);
  std::fputc('\n', f);
  std::fflush(f);
}
}  // namespace

void log_set_fp(FILE *fp) { g_fp = fp; }

void log_log(int level, const char *file, int line, const char *fmt, ...) {
  using namespace std::chrono;
  const auto now = system_clock::now();
  const double ts = duration_cast<nanoseconds>(now.time_since_epoch()).count() / 1e9;
  const double et = duration_cast<nanoseconds>(steady_clock::now() - g_start).count() / 1e9;
  const std::time_t t = system_clock::to_time_t(now);
  std::tm lt;
  localtime_r(&t, &lt);
  char stamp[32];
  stamp[std::strftime(stamp, sizeof(stamp), "%Y-%m-%d %H:%M:%S", &lt)] = '\0';
  const char *base = std::strrchr(file, '/');
  base = base ? base + 1 : file;
  std::lock_guard<std::mutex> lock(g_mu);
  va_list ap;
  va_start(ap, fmt);
  emit(stderr, stamp, level, ts, et, base, line, fmt, ap);
  va_end(ap);
  if (g_fp) {
    va_start(ap, fmt);
    emit(g_fp, stamp, level, ts, et, base, line, fmt, ap);
    va_end(ap);
  }
}

std::string FormatWithCommas(long long value) {
  std::stringstream ss;
  try {
    ss.imbue(std::locale(""));
  } catch (...) {
  }
  ss << std::fixed << value;
  return ss.str();
}

long getValue() {
  FILE *f = std::fopen("/proc/self/status", "r");
  if (!f) return -1;
  long kb = -1;
  char line[256];
  while (std::fgets(line, sizeof(line), f))
    if (std::strncmp(line, "VmRSS:", 6) == 0) {
      kb = std::strtol(line + 6, nullptr, 10);
      break;
    }
  std::fclose(f);
  return kb;
}
