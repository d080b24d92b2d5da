// Logging macros of the Blade C++ API (reference include/blade/logger.hh:81-143).
// Messages use "{}" placeholders, filled in order by operator<<.  BL_TRACE
// prints only when the environment variable TRACE is set; BL_DEBUG only when
// BLADE_DEBUG is set; BLADE_QUIET silences INFO.
#ifndef BLADE_LOGGER_HH
#define BLADE_LOGGER_HH

#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>

namespace Blade::Log {

inline void format_into(std::ostringstream& os, const char* fmt) { os << fmt; }

template <typename T, typename... Rest>
void format_into(std::ostringstream& os, const char* fmt, const T& v, const Rest&... rest) {
    for (const char* p = fmt; *p; ++p) {
        if (*p == '{') {
            const char* q = p;
            while (*q && *q != '}') ++q;
            if (*q == '}') {
                os << v;
                format_into(os, q + 1, rest...);
                return;
            }
        }
        os << *p;
    }
}

template <typename... Args>
std::string format(const char* fmt, const Args&... args) {
    std::ostringstream os;
    format_into(os, fmt, args...);
    return os.str();
}

inline bool enabled(const char* env) { return std::getenv(env) != nullptr; }

template <typename... Args>
void emit(FILE* f, const char* level, const char* domain, const char* fmt, const Args&... args) {
    const std::string msg = format(fmt, args...);
    std::fprintf(f, "BLADE [%s] [%s] %s\n", level, domain, msg.c_str());
}

}  // namespace Blade::Log

#ifndef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "BLADE"
#endif

#define BL_TRACE(...) \
    do { if (::Blade::Log::enabled("TRACE")) ::Blade::Log::emit(stdout, "TRACE", BL_LOG_DOMAIN, __VA_ARGS__); } while (0)
#define BL_DEBUG(...) \
    do { if (::Blade::Log::enabled("BLADE_DEBUG")) ::Blade::Log::emit(stdout, "DEBUG", BL_LOG_DOMAIN, __VA_ARGS__); } while (0)
#define BL_INFO(...) \
    do { if (!::Blade::Log::enabled("BLADE_QUIET")) ::Blade::Log::emit(stdout, "INFO", BL_LOG_DOMAIN, __VA_ARGS__); } while (0)
#define BL_WARN(...) ::Blade::Log::emit(stderr, "WARN", BL_LOG_DOMAIN, __VA_ARGS__)
#define BL_ERROR(...) ::Blade::Log::emit(stderr, "ERROR", BL_LOG_DOMAIN, __VA_ARGS__)
#define BL_FATAL(...) ::Blade::Log::emit(stderr, "FATAL", BL_LOG_DOMAIN, __VA_ARGS__)

#endif
