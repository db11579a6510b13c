// Logging for the drop-in executables.  Lines have the reference's layout
// (scan-query/util/log.cpp:101-150):
//   YYYY-mm-dd HH:MM:SS LEVEL (ts: <epoch s>, et: <elapsed s>) file:line: message
// so the reference's log parsers (python_experiments/data_analysis/util/read_file_utils.py)
// read them unchanged.
#pragma once

#include <cstdio>
#include <string>

enum { ACC_LOG_TRACE, ACC_LOG_DEBUG, ACC_LOG_INFO, ACC_LOG_WARN, ACC_LOG_ERROR, ACC_LOG_FATAL };

void log_set_fp(FILE *fp);
void log_log(int level, const char *file, int line, const char *fmt, ...) __attribute__((format(printf, 4, 5)));

#define log_trace(...) log_log(ACC_LOG_TRACE, __FILE__, __LINE__, __VA_ARGS__)
#define log_debug(...) log_log(ACC_LOG_DEBUG, __FILE__, __LINE__, __VA_ARGS__)
#define log_info(...) log_log(ACC_LOG_INFO, __FILE__, __LINE__, __VA_ARGS__)
#define log_warn(...) log_log(ACC_LOG_WARN, __FILE__, __LINE__, __VA_ARGS__)
#define log_error(...) log_log(ACC_LOG_ERROR, __FILE__, __LINE__, __VA_ARGS__)

// Digit grouping of the environment's locale, as FormatWithCommas (scan-query/util/util.h:17-22).
std::string FormatWithCommas(long long value);
// Resident set size in KB from /proc/self/status (scan-query/util/stat.cpp:13-25).
long getValue();
